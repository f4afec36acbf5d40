"""The FMM's treecode and outlier modes (shp_fmm_set_type) on the GPU against
the oracle's FP64 restatement of the same modes on the same tree (leaf level
fixed), and against the FP64 direct sum.  Bar: 1e-5 relative L2 (global
norm) against the oracle and against the direct sum, the plain FMM's bar for
FP32 expansions, M2P and pairs; the outlier flags exactly."""
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
    e.fmm_set_type(solveit_hip.SHP_FMM_FMM)
    e.close()


def _rel(a, b, rows=None):
    if rows is not None:
        a, b = a[rows], b[rows]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _plume_with_outliers(n, seed, far=5e-5):
    d = synth.cone(n, seed=seed)
    rng = np.random.default_rng(seed)
    idx = rng.choice(n, 13, replace=False)
    d["x"][idx[:4]] += far * (1 + rng.random(4))
    d["y"][idx[4:8]] -= far * (1 + rng.random(4))
    d["z"][idx[8:12]] += 3 * far * (1 + rng.random(4))
    d["x"][idx[12]] = 1e-2          # one very far ion: the whole tree is one expansion for it
    d["status"][idx[1::3]] = 2      # some outliers are sources only
    d["status"][rng.choice(n, 40, replace=False)] = 0
    return d


def _gpu(eng, d, fmm_type, p, L):
    import solveit_hip
    import torch
    eng.fmm_set_type(fmm_type)
    try:
        st = solveit_hip.DeviceState.from_numpy(d)
        a = eng.fmm_dev(st, order=p, leaf_level=L, ring=2)
        return torch.stack(a, 1).cpu().numpy()
    finally:
        eng.fmm_set_type(solveit_hip.SHP_FMM_FMM)


def test_outlier_flags_equal_oracle(eng):
    import solveit_hip
    d = _plume_with_outliers(6000, 41)
    print("case outlier flags, 6000-ion cone with planted outliers", flush=True)
    flags, k = eng.fmm_outliers(solveit_hip.DeviceState.from_numpy(d))
    ref, kr = oracle.fmm_outliers(d)
    np.testing.assert_array_equal(flags, ref)
    assert k == kr > 13


@pytest.mark.parametrize("p", [8, 10])
def test_treecode_matches_oracle(eng, p):
    import solveit_hip
    d = synth.cone(4000, seed=42)
    print(f"case treecode p={p}, 4000-ion cone, leaf level 4", flush=True)
    a = _gpu(eng, d, solveit_hip.SHP_FMM_TREECODE, p, 4)
    ref, st = oracle.fmm_f64(d, p, 4, ring=2, fmm_type=oracle.FMM_TREECODE)
    assert st["m2p"] > 0
    direct = np.asarray(oracle.direct_f64(d))
    tgt = d["status"] == 1
    print(f"  gpu vs oracle {_rel(a, ref, tgt):.2e}, vs direct {_rel(a, direct, tgt):.2e}", flush=True)
    assert _rel(a, ref, tgt) < 1e-5
    assert _rel(a, direct, tgt) < 1e-5
    assert np.all(a[~tgt] == 0)


@pytest.mark.parametrize("kind", ["fmm", "treecode"])
def test_outlier_modes_match_oracle(eng, kind):
    import solveit_hip
    d = _plume_with_outliers(5000, 43)
    print(f"case {kind} with outliers, p=10, leaf level 4", flush=True)
    gt = solveit_hip.SHP_FMM_FMM_OUTLIERS if kind == "fmm" else solveit_hip.SHP_FMM_TREECODE_OUTLIERS
    ot = oracle.FMM_FMM_OUTLIERS if kind == "fmm" else oracle.FMM_TREECODE_OUTLIERS
    a = _gpu(eng, d, gt, 10, 4)
    oracle.fmm_set_shell_orders([0, 0, 0] + list(solveit_hip.fmm_shell_orders(10)))
    try:
        ref, _ = oracle.fmm_f64(d, 10, 4, ring=2, fmm_type=ot)
    finally:
        oracle.fmm_set_shell_orders(None)
    direct = np.asarray(oracle.direct_f64(d))
    tgt = d["status"] == 1
    flags, _ = oracle.fmm_outliers(d)
    rows = (flags == 1) & tgt
    print(f"  gpu vs oracle {_rel(a, ref, tgt):.2e} (outliers {_rel(a, ref, rows):.2e}), "
          f"vs direct {_rel(a, direct, tgt):.2e}", flush=True)
    assert rows.sum() >= 5
    assert _rel(a, ref, tgt) < 1e-5
    assert _rel(a, ref, rows) < 1e-5
    assert _rel(a, direct, tgt) < 1e-5
    assert np.all(a[~tgt] == 0)


def test_mode_limits(eng):
    import solveit_hip
    d = synth.cone(2000, seed=44)
    st = solveit_hip.DeviceState.from_numpy(d)
    eng.fmm_set_type(solveit_hip.SHP_FMM_TREECODE)
    try:
        with pytest.raises(solveit_hip.ShpError):
            eng.fmm_dev(st, order=12, leaf_level=3, ring=2)
        eng.fmm_set_type(solveit_hip.SHP_FMM_FMM_OUTLIERS)
        with pytest.raises(solveit_hip.ShpError):
            eng.fmm_dev(st, order=8, leaf_level=3, ring=2, symmetry=solveit_hip.make_symmetry("rotational", nsec=4))
    finally:
        eng.fmm_set_type(solveit_hip.SHP_FMM_FMM)
    with pytest.raises(solveit_hip.ShpError):
        eng.fmm_set_type(7)
