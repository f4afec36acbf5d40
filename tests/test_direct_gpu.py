"""Parity of the HIP direct sums with the oracle, through the C ABI.

Bars: REF and FP64 contracts within 1e-12 relative per ion (double sums in a
different order); FP32C within rel-L2 1e-5 of the FP64 oracle (the tolerance
the force engine is specified to)."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12
TOL_FP32C_L2 = 1e-5


@pytest.fixture(scope="module")
def eng():
    import solveit_hip
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    yield e
    e.close()


def _rowerr(a, ref):
    na = np.linalg.norm(ref, axis=1)
    e = np.linalg.norm(a - ref, axis=1)
    return np.where(na > 0, e / np.where(na > 0, na, 1), e)


def _dev(d):
    import solveit_hip
    import torch
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    return st


def _direct_dev(eng, d, prec, sym=None):
    import torch
    st = _dev(d)
    ax, ay, az = eng.direct_dev(st, precision=prec, symmetry=sym)
    eng.sync()
    return torch.stack([ax, ay, az], 1).cpu().numpy()


@pytest.mark.parametrize("n", [1, 2, 255, 257, 1000, 2048])
def test_ref_plain_hemisphere(eng, n):
    d = synth.hemisphere(n, seed=100 + n)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    print(f"case ref-plain n={n}", flush=True)
    got = eng.direct_host(pos4, vel4, acc4)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    assert _rowerr(got, ref).max() <= TOL_FP64
    # the adapter's float rounding then matches the reference's bit for bit
    np.testing.assert_array_equal(got.astype(np.float32), ref.astype(np.float32))


def test_ref_ranges_and_status(eng):
    d = synth.cone(3000, seed=21)
    d["status"][::7] = 2
    d["status"][::11] = 0
    d["status"][::13] = -1
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    print("case ref-ranges", flush=True)
    for (i0, i1, j0, j1) in [(0, 3000, 0, 3000), (100, 1900, 0, 3000), (0, 3000, 500, 2500),
                             (5, 6, 0, 3000), (0, 3000, 7, 8), (10, 10, 0, 3000)]:
        got = eng.direct_host(pos4, vel4, acc4, i0, i1, j0, j1)
        ref = oracle.direct_ref(pos4, vel4, acc4, i0, i1, j0, j1)
        assert got.shape == ref.shape
        if got.size:
            assert _rowerr(got, ref).max() <= TOL_FP64


@pytest.mark.parametrize("nsec", [4, 6])
def test_ref_rotational(eng, nsec):
    d = synth.cone(1500, seed=31)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    import solveit_hip
    print(f"case ref-rot nsec={nsec}", flush=True)
    sym = solveit_hip.make_symmetry("rotational", nsec=nsec)
    got = eng.direct_host(pos4, vel4, acc4, symmetry=sym)
    ref = oracle.direct_ref(pos4, vel4, acc4, sym="rotational", nsec=nsec)
    assert _rowerr(got, ref).max() <= TOL_FP64


def test_ref_translational(eng):
    import solveit_hip
    d = synth.cell_fill(1200, seed=41)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    lat = oracle.reference_lattice(3, (1, 1, 1))
    dom = (2e-6, 2e-6, 2e-6)
    print("case ref-trans", flush=True)
    sym = solveit_hip.make_symmetry("translational", lattice=lat, domain=dom)
    got = eng.direct_host(pos4, vel4, acc4, symmetry=sym)
    ref = oracle.direct_ref(pos4, vel4, acc4, sym="translational", lattice=lat, domain=dom)
    assert _rowerr(got, ref).max() <= TOL_FP64


@pytest.mark.parametrize("kind", ["none", "rotational", "translational"])
def test_fp64_dev(eng, kind):
    import solveit_hip
    d = synth.cone(2500, seed=51) if kind != "translational" else synth.cell_fill(1500, seed=52)
    d["status"][::17] = 2
    d["status"][::19] = 0
    lat = synth.full_lattice(1, (1, 1, 1))
    dom = (2e-6, 2e-6, 2e-6)
    sym = solveit_hip.make_symmetry(kind, nsec=6, lattice=lat, domain=dom)
    print(f"case fp64-dev {kind}", flush=True)
    got = _direct_dev(eng, d, solveit_hip.SHP_PREC_FP64, sym)
    ref = oracle.direct_f64(d, sym=kind, nsec=6, lattice=lat, domain=dom)
    assert _rowerr(got, ref).max() <= TOL_FP64
    assert np.all(got[d["status"] != 1] == 0)


@pytest.mark.parametrize("kind", ["none", "rotational", "translational"])
def test_fp32c_dev_vs_fp64_oracle(eng, kind):
    import solveit_hip
    d = synth.cone(12000, seed=61) if kind != "translational" else synth.cell_fill(3000, seed=62)
    lat = synth.full_lattice(1, (1, 1, 1))
    dom = (2e-6, 2e-6, 2e-6)
    sym = solveit_hip.make_symmetry(kind, nsec=6, lattice=lat, domain=dom)
    print(f"case fp32c-dev {kind}", flush=True)
    got = _direct_dev(eng, d, solveit_hip.SHP_PREC_FP32C, sym)
    ref = oracle.direct_f64(d, sym=kind, nsec=6, lattice=lat, domain=dom)
    l2, mx, _ = oracle.rel_l2(got, ref, d["status"] == 1)
    print(f"  fp32c {kind}: rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= TOL_FP32C_L2


def test_fp32c_vs_fp64_gpu_full_size(eng):
    """FP32C against the GPU FP64 path on 262,144 cone ions."""
    import solveit_hip
    d = synth.cone(262144)
    print("case fp32c-vs-fp64 n=262144", flush=True)
    a32 = _direct_dev(eng, d, solveit_hip.SHP_PREC_FP32C)
    a64 = _direct_dev(eng, d, solveit_hip.SHP_PREC_FP64)
    l2, mx, _ = oracle.rel_l2(a32, a64)
    print(f"  fp32c vs fp64 at 262144: rel L2 {l2:.3e} max {mx:.3e}")
    assert l2 <= TOL_FP32C_L2
    # momentum identity on the FP64 path at full size
    m = np.array([oracle.species_mass(s) for s in d["species"]])
    p = (a64 * m[:, None]).sum(0)
    assert np.all(np.abs(p) <= 1e-10 * np.abs(a64 * m[:, None]).sum(0))


def test_fp32c_host_matches_ref(eng):
    import solveit_hip
    d = synth.cone(4000, seed=71)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    print("case fp32c-host", flush=True)
    got = eng.direct_host(pos4, vel4, acc4, precision=solveit_hip.SHP_PREC_FP32C)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    assert oracle.rel_l2(got, ref)[0] <= TOL_FP32C_L2


def test_bad_arguments_raise(eng):
    import solveit_hip
    d = synth.hemisphere(10)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    with pytest.raises(solveit_hip.ShpError):
        eng.direct_host(pos4, vel4, acc4, 0, 11, 0, 10)
    with pytest.raises(solveit_hip.ShpError):
        eng.direct_host(pos4, vel4, acc4, precision=solveit_hip.SHP_PREC_FP64)
