"""Pins the oracle's treecode and outlier FMM modes (CPU only).

* Outlier flags equal the statistic computed independently with numpy
  (median, median absolute deviation x 1.4826, threshold 3, valid ions only).
* Treecode (M2P from every level's far boxes, no local expansions) converges
  to the FP64 direct sum as the order grows, and a far cluster's field by M2P
  matches the direct sum to the expansion's truncation.
* The outlier modes match the direct sum over all ions, with the outliers
  held out of the tree; without outliers they equal the plain modes bit for
  bit.
"""
import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _rel(a, b, rows=None):
    if rows is not None:
        a, b = a[rows], b[rows]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _direct(d):
    return np.asarray(oracle.direct_f64(d))


def _with_outliers(n, seed, far=5e-5):
    d = synth.cone(n, seed=seed)
    rng = np.random.default_rng(seed)
    k = 12
    idx = rng.choice(n, k, replace=False)
    d["x"][idx[:4]] += far * (1 + rng.random(4))
    d["y"][idx[4:8]] -= far * (1 + rng.random(4))
    d["z"][idx[8:]] += 3 * far * (1 + rng.random(4))
    d["status"][idx[::3]] = 2   # some outliers are sources only
    d["status"][rng.choice(n, 40, replace=False)] = 0
    return d, idx


def test_outlier_flags_equal_the_numpy_statistic():
    d, planted = _with_outliers(3000, 31)
    flags, k = oracle.fmm_outliers(d)
    valid = (d["status"] == 1) | (d["status"] == 2)
    want = np.zeros(len(valid), bool)
    for c in ("x", "y", "z"):
        v = d[c][valid]
        med = np.median(v)
        mad = float(np.float32(1.482602218505602)) * np.median(np.abs(v - med))
        want[valid] |= np.abs(v - med) / mad >= 3.0
    np.testing.assert_array_equal(flags.astype(bool), want)
    assert k == want.sum()
    assert all(flags[i] for i in planted if valid[i])


@pytest.mark.parametrize("p", [4, 6, 8, 10])
def test_treecode_converges_to_direct(p):
    d = synth.cone(2500, seed=32)
    ref = _direct(d)
    a, st = oracle.fmm_f64(d, p, 4, ring=2, fmm_type=oracle.FMM_TREECODE)
    assert st["m2p"] > 0 and st["m2l"] == 0
    err = _rel(a, ref)
    bound = {4: 3e-5, 6: 1e-6, 8: 1e-7, 10: 5e-9}[p]
    assert err < bound, err


def test_treecode_far_cluster_matches_direct():
    rng = np.random.default_rng(33)
    a = rng.uniform(0, 1e-7, (300, 3))
    b = rng.uniform(0, 1e-7, (300, 3)) + np.array([1e-5, 0.0, 0.0])
    pos = np.vstack([a, b])
    d = synth._pack(pos, np.where(rng.random(600) < 0.1, -1.0, 1.0), np.zeros(600, np.int32),
                    np.zeros((600, 3)))
    ref = _direct(d)
    errs = []
    for p in (2, 4, 6, 8):
        out, st = oracle.fmm_f64(d, p, 4, ring=1, fmm_type=oracle.FMM_TREECODE)
        assert st["m2p"] > 0
        errs.append(_rel(out, ref))
    assert errs[-1] < 1e-6
    assert all(e1 < e0 for e0, e1 in zip(errs, errs[1:]))


@pytest.mark.parametrize("fmm_type", ["fmm", "treecode"])
def test_outlier_modes_match_direct(fmm_type):
    d, planted = _with_outliers(2500, 34)
    ref = _direct(d)
    t = oracle.FMM_FMM_OUTLIERS if fmm_type == "fmm" else oracle.FMM_TREECODE_OUTLIERS
    a, st = oracle.fmm_f64(d, 10, 4, ring=2, fmm_type=t)
    tgt = d["status"] == 1
    assert _rel(a, ref, tgt) < 3e-6
    flags, _ = oracle.fmm_outliers(d)
    rows = (flags == 1) & tgt
    assert rows.sum() >= 5
    assert _rel(a, ref, rows) < 1e-7  # the outliers' own accelerations
    assert st["m2p"] > 0
    assert np.all(a[~tgt] == 0)


def test_outlier_modes_without_outliers_equal_plain_modes():
    rng = np.random.default_rng(35)
    d = synth._pack(rng.uniform(-1e-6, 1e-6, (1500, 3)), np.ones(1500), np.zeros(1500, np.int32),
                    np.zeros((1500, 3)))
    flags, k = oracle.fmm_outliers(d)
    assert k == 0
    for plain, outl in ((oracle.FMM_FMM, oracle.FMM_FMM_OUTLIERS),
                        (oracle.FMM_TREECODE, oracle.FMM_TREECODE_OUTLIERS)):
        a0, _ = oracle.fmm_f64(d, 6, 3, ring=2, fmm_type=plain)
        a1, _ = oracle.fmm_f64(d, 6, 3, ring=2, fmm_type=outl)
        np.testing.assert_array_equal(a0, a1)
