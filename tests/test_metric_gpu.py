"""The accuracy metric of Integrator::compareAcceleration
(integration_schemes.cpp:374-470) on the GPU (shp_compare_accel) against the
oracle's restatement (oracle.rel_l2) and hand cases: status masks, rows where
both the error and the reference vanish, ties of the maximum (first index),
no valid ions (NaN, as the reference's 0/0)."""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    import solveit_hip
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    yield e
    e.close()


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda:0")


def _cmp(eng, status, a, ref):
    import torch
    return eng.compare_accel(_t(status, torch.int32), tuple(_t(a[:, k]) for k in range(3)),
                             tuple(_t(ref[:, k]) for k in range(3)))


@pytest.mark.parametrize("n", [1, 255, 256, 70001, 1 << 20])
def test_metric_matches_oracle(eng, n):
    rng = np.random.default_rng(n)
    ref = rng.normal(size=(n, 3))
    a = ref * (1 + 1e-5 * rng.normal(size=(n, 3)))
    status = np.where(rng.random(n) < 0.9, 1, rng.choice([0, 2, -1], n)).astype(np.int32)
    status[0] = 1
    if n > 10:
        ref[3] = 0.0          # both zero: counts 0
        a[3] = 0.0
        status[3] = 1
    print(f"case metric n={n}", flush=True)
    got = _cmp(eng, status, a, ref)
    mask = status == 1
    l2, mx, k = oracle.rel_l2(a, ref, mask)
    assert got["n"] == int(mask.sum())
    assert abs(got["l2"] - l2) <= 1e-12 * l2
    assert abs(got["max"] - mx) <= 1e-14 * mx
    assert got["idx"] == int(np.nonzero(mask)[0][k])


def test_metric_hand_cases(eng):
    ref = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 0], [3.0, 0, 0], [0, 0, 1.0]])
    a = np.array([[1.1, 0, 0], [0, 2.2, 0], [0, 0, 0], [3.0, 0, 0], [0, 0, 1.1]])
    st = np.array([1, 1, 1, 1, 0], np.int32)
    got = _cmp(eng, st, a, ref)
    # rel = 0.01, 0.01, 0 (both zero), 0; ion 4 not inbound; first maximum is ion 0
    assert got["n"] == 4
    assert abs(got["l2"] - math.sqrt(0.02 / 4)) <= 1e-15
    assert abs(got["max"] - 0.1) <= 1e-15
    assert got["idx"] == 0
    st0 = np.zeros(5, np.int32)
    got0 = _cmp(eng, st0, a, ref)
    assert got0["n"] == 0 and math.isnan(got0["l2"]) and got0["max"] == 0.0 and got0["idx"] == 0
