"""Parity of the HIP FMM with the oracle's FP64 FMM (same tree, order and
near-field width) and accuracy against the FP64 direct sum, through the C ABI.

Bars: FP32 GPU FMM vs FP64 oracle FMM within rel-L2 1e-5 (expansion rounding);
GPU FMM vs FP64 direct within rel-L2 1e-5 at p=10 (the engine's accuracy
target); at N = 1,048,576 against the GPU FP64 direct path on a seeded
8192-ion subsample."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import solveit_hip
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    yield e
    e.close()


def _fmm(eng, d, **kw):
    import solveit_hip
    import torch
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    out = eng.fmm_dev(st, **kw)
    eng.sync()
    return torch.stack(out, 1).cpu().numpy(), eng.fmm_leaf_level()


def _oracle_fmm(d, p, L, ring):
    """Oracle FMM with the engine's M2L order per far-band shell (the engine
    reports them through shp_fmm_shell_orders)."""
    import solveit_hip
    orders = solveit_hip.fmm_shell_orders(p, ring)
    if orders[0]:
        oracle.fmm_set_shell_orders([0, 0, 0] + list(orders))
    try:
        return oracle.fmm_f64(d, p, L, ring=ring)
    finally:
        oracle.fmm_set_shell_orders(None)


@pytest.mark.parametrize("p", [4, 6, 8, 10])
def test_fmm_matches_oracle_fmm(eng, p):
    d = synth.cone(8192, seed=21)
    d["status"][::23] = 2
    d["status"][::29] = 0
    print(f"case fmm-vs-oracle p={p}", flush=True)
    got, L = _fmm(eng, d, order=p, leaf_level=5, ring=2)
    ref, st = _oracle_fmm(d, p, L, 2)
    assert L == 5 and st["m2l"] > 0
    l2, mx, _ = oracle.rel_l2(got, ref, d["status"] == 1)
    print(f"  p={p}: gpu fmm vs oracle fmm rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5
    assert np.all(got[d["status"] != 1] == 0)


def test_fmm_ring1_matches_oracle(eng):
    d = synth.cone(6000, seed=22)
    print("case fmm ring1", flush=True)
    got, L = _fmm(eng, d, order=6, leaf_level=5, ring=1)
    ref, _ = _oracle_fmm(d, 6, L, 1)
    assert oracle.rel_l2(got, ref)[0] <= 1e-5


def test_fmm_accuracy_vs_direct(eng):
    d = synth.cone(16384, seed=23)
    print("case fmm-vs-direct", flush=True)
    got, L = _fmm(eng, d, order=10, leaf_level=0, leaf_size=20, ring=2)
    ref = oracle.direct_f64(d)
    l2, mx, _ = oracle.rel_l2(got, ref)
    print(f"  leaf level {L}: fmm p=10 vs fp64 direct rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5


def test_fmm_edge_cases(eng):
    print("case fmm edge", flush=True)
    d = synth.hemisphere(1, seed=1)
    got, _ = _fmm(eng, d, order=4, leaf_level=3, ring=2)
    assert np.all(got == 0)
    d = synth.hemisphere(2, seed=2)
    got, _ = _fmm(eng, d, order=4, leaf_level=1, ring=2)
    ref = oracle.direct_f64(d)
    # two ions in the near field of each other: FP32 pair arithmetic only
    assert oracle.rel_l2(got, ref)[0] < 1e-6
    d = synth.cone(300, seed=3)
    d["status"][:] = 0
    got, _ = _fmm(eng, d, order=4, leaf_level=3, ring=2)
    assert np.all(got == 0)


def test_fmm_coincident_ions(eng):
    """Clusters of ions at one point each (identical Morton keys at every
    level) next to a scattered plume, automatic leaf level: coincident pairs
    contribute nothing (zero displacement, softened distance), so the FMM
    still matches the FP64 direct sum; a plume that is one point throughout
    gives zero forces at every automatic depth."""
    d = synth.cone(4000, seed=31)
    for k, (a, b) in enumerate(((0, 600), (600, 700), (700, 701))):
        j = 1000 + 7 * k
        for c in ("x", "y", "z"):
            d[c][a:b] = d[c][j]
    print("case fmm coincident", flush=True)
    got, L = _fmm(eng, d, order=10, leaf_level=0, leaf_size=20, ring=2)
    ref = oracle.direct_f64(d)
    l2, mx, _ = oracle.rel_l2(got, ref)
    print(f"  leaf level {L}: rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5
    one = synth.cone(500, seed=32)
    for c in ("x", "y", "z"):
        one[c][:] = one[c][0]
    got, L = _fmm(eng, one, order=10, leaf_level=0, leaf_size=20, ring=2)
    print(f"  single point: leaf level {L}")
    assert np.all(got == 0)


def test_fmm_full_size_vs_fp64_direct(eng):
    import solveit_hip
    import torch
    n = 1048576
    d = synth.cone(n)
    print("case fmm n=1048576", flush=True)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    a_f = torch.stack(eng.fmm_dev(st, order=10, leaf_level=0, leaf_size=20, ring=2), 1)
    L = eng.fmm_leaf_level()
    a_d = torch.stack(eng.direct_dev(st, precision=solveit_hip.SHP_PREC_FP64), 1)
    sub = np.random.default_rng(8192).choice(n, 8192, replace=False)
    l2, mx, _ = oracle.rel_l2(a_f.cpu().numpy()[sub], a_d.cpu().numpy()[sub])
    print(f"  n=1M leaf level {L}: fmm p=10 vs gpu fp64 direct rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5


@pytest.mark.parametrize("n,L", [(65536, 7), (1048576, 8)])
def test_fmm_tree_counts_match_oracle(eng, n, L):
    """The GPU tree's interaction lists equal the tree definition (exact counts)."""
    import solveit_hip
    import torch
    d = synth.cone(n)
    print(f"case fmm counts n={n} L={L}", flush=True)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    eng.fmm_dev(st, order=2, leaf_level=L, ring=2)
    s = eng.fmm_stats()
    ref = oracle.fmm_counts(d, L, ring=2)
    print(f"  gpu m2l {s['m2l_pairs']:.0f} p2p {s['p2p_pairs']:.0f}; "
          f"oracle m2l {ref['m2l']:.0f} p2p {ref['p2p_pairs']:.0f}")
    assert s["leaf_level"] == L
    assert s["m2l_pairs"] == ref["m2l"]
    assert s["p2p_pairs"] == ref["p2p_pairs"]


@pytest.mark.parametrize("sym", ["rotational", "translational"])
def test_fmm_symmetry_vs_fp64_direct_images(eng, sym):
    """Rotational (6 sectors) and translational (3x3x3 shell of 2 um cells)
    FMM against the FP64 direct sum over the same explicit images."""
    import solveit_hip
    import torch
    if sym == "rotational":
        d = synth.cone(20000, seed=24)
        s = solveit_hip.make_symmetry("rotational", nsec=6)
    else:
        d = synth.cell_fill(12000, seed=25)
        s = solveit_hip.make_symmetry("translational", lattice=synth.full_lattice(1, (1, 1, 1)),
                                      domain=(2e-6, 2e-6, 2e-6))
    print(f"case fmm {sym}", flush=True)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    got = torch.stack(eng.fmm_dev(st, order=10, leaf_size=20, ring=2, symmetry=s), 1).cpu().numpy()
    ref = torch.stack(eng.direct_dev(st, precision=solveit_hip.SHP_PREC_FP64, symmetry=s), 1)
    l2, mx, _ = oracle.rel_l2(got, ref.cpu().numpy(), d["status"] == 1)
    print(f"  {sym}: fmm p=10 vs fp64 direct over images rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5


def test_leaf_level_choice_counts(eng):
    """The leaf-level choice counts, on every box of each candidate level (no
    sampling below 4096 boxes), the near-field particle pairs and far-band
    box pairs; for the chosen level they equal the pairs the leaf and M2L
    kernels then evaluate, and the chosen level has the least estimate."""
    import solveit_hip
    import torch
    d = synth.cone(20000, seed=26)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    print("case fmm level choice", flush=True)
    for leaf_size in (200, 60):
        eng.fmm_dev(st, order=10, leaf_level=0, leaf_size=leaf_size, ring=2)
        rows = eng.fmm_level_estimates()
        s = eng.fmm_stats()
        L = eng.fmm_leaf_level()
        print(f"  leaf_size {leaf_size}: chose {L}; " +
              ", ".join(f"L{r['level']}: p2p {r['p2p_pairs']:.3e} far {r['far_pairs']:.3e} "
                        f"{r['est_ms']:.3f} ms" for r in rows))
        assert rows and s["leaf_level"] == L
        best = min(rows, key=lambda r: r["est_ms"])
        assert best["level"] == L
        chosen = [r for r in rows if r["level"] == L][0]
        assert chosen["p2p_pairs"] == s["p2p_pairs"]
        if rows[0]["level"] == 2:  # every M2L level is a candidate
            assert sum(r["far_pairs"] for r in rows if r["level"] <= L) == s["m2l_pairs"]
    eng.fmm_dev(st, order=10, leaf_level=5, ring=2)
    assert eng.fmm_level_estimates() == []


def _float_state(d):
    """The plume with float-representable positions and charges, so that the
    host AoS entry (float positions, as the reference stores them) and the
    FP64 oracle see the same ions."""
    for k in ("x", "y", "z", "q"):
        d[k] = d[k].astype(np.float32).astype(np.float64)
    return d


@pytest.mark.parametrize("p", [6, 10])
def test_fmm_host_matches_oracle_fmm(eng, p):
    """shp_fmm_host (the reference's vec4<float> AoS, [first, last) ranges)
    against the oracle's FP64 FMM of the same tree: rel L2 <= 1e-5."""
    d = _float_state(synth.cone(8192, seed=24))
    d["status"][::23] = 2
    d["status"][::29] = 0
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    print(f"case fmm host p={p}", flush=True)
    got = eng.fmm_host(pos4, vel4, acc4, order=p, leaf_level=5, ring=2)
    ref, _ = _oracle_fmm(d, p, 5, 2)
    l2, mx, _ = oracle.rel_l2(got, ref, d["status"] == 1)
    print(f"  p={p}: host fmm vs oracle fmm rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= 1e-5
    assert np.all(got[d["status"] != 1] == 0)
    # the host arrays are left as they were
    p2, v2, a2 = synth.to_reference_aos(d)
    assert np.array_equal(pos4, p2) and np.array_equal(vel4, v2) and np.array_equal(acc4, a2)


def test_fmm_host_subrange_and_symmetry(eng):
    """[i0, i1) evaluates the sub-plume alone (the ions outside the range are
    neither sources nor targets), and the symmetric variant equals
    shp_fmm_sym_dev on the same ions."""
    import solveit_hip
    import torch
    d = _float_state(synth.cone(6000, seed=25))
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    print("case fmm host subrange", flush=True)
    got = eng.fmm_host(pos4, vel4, acc4, i0=1000, i1=5000, order=8, leaf_level=4, ring=2)
    sub = {k: v[1000:5000].copy() for k, v in d.items()}
    ref, _ = _oracle_fmm(sub, 8, 4, 2)
    assert oracle.rel_l2(got, ref)[0] <= 1e-5
    print("case fmm host rotational", flush=True)
    s = solveit_hip.make_symmetry("rotational", nsec=6)
    got = eng.fmm_host(pos4, vel4, acc4, order=10, leaf_level=4, ring=2, symmetry=s)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    dev = torch.stack(eng.fmm_dev(st, order=10, leaf_level=4, ring=2, symmetry=s), 1).cpu().numpy()
    np.testing.assert_array_equal(got, dev)
