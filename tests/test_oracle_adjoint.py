"""Pins the oracle's force derivatives and discrete adjoint (CPU only):
the JVP against central differences of the FP64 direct sum, the VJP against
the JVP through the dot-product identity <b, J d> = <J^T b, d>, and the
rollout gradient of a plume-divergence objective against central differences
of the rollout itself."""
import numpy as np

import oracle
from solveit_hip import synth


def _small(n=200, seed=61):
    d = synth.hemisphere(n, seed=seed)
    d["status"][::11] = 2
    d["q"][::7] = -1.0
    d["species"][::7] = 5
    return d


def test_force_jvp_matches_central_differences():
    d = _small()
    rng = np.random.default_rng(1)
    n = d["x"].shape[0]
    dx, dy, dz = (rng.normal(0, 1e-9, n) for _ in range(3))
    dq = rng.normal(0, 0.1, n)
    jvp = oracle.force_jvp_f64(d, dx, dy, dz, dq)
    h = 1e-4
    dp, dm = dict(d), dict(d)
    for k, t in (("x", dx), ("y", dy), ("z", dz), ("q", dq)):
        dp[k] = d[k] + h * t
        dm[k] = d[k] - h * t
    fd = (oracle.direct_f64(dp) - oracle.direct_f64(dm)) / (2 * h)
    assert np.linalg.norm(jvp - fd) <= 1e-6 * np.linalg.norm(fd)


def test_force_vjp_dot_product_identity():
    d = _small()
    rng = np.random.default_rng(2)
    n = d["x"].shape[0]
    dx, dy, dz = (rng.normal(0, 1e-9, n) for _ in range(3))
    dq = rng.normal(0, 0.1, n)
    b = rng.normal(0, 1, (n, 3))
    lhs = np.sum(b * oracle.force_jvp_f64(d, dx, dy, dz, dq))
    xb, qb = oracle.force_vjp_f64(d, b)
    rhs = np.sum(xb[:, 0] * dx + xb[:, 1] * dy + xb[:, 2] * dz) + np.sum(qb * dq)
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


def _divergence(x, v):
    """Plume divergence objective: mean squared radial distance; its gradient."""
    n = x.shape[0]
    j = np.sum(x[:, 0] ** 2 + x[:, 1] ** 2) / n
    g = np.zeros_like(x)
    g[:, 0] = 2 * x[:, 0] / n
    g[:, 1] = 2 * x[:, 1] / n
    return j, g, np.zeros_like(v)


def _objective(d, dt, steps):
    _, _, _, fin = oracle.rollout_grad_f64(d, dt, steps, lambda x, v: (x * 0, v * 0))
    x = np.stack([fin["x"], fin["y"], fin["z"]], 1)
    return _divergence(x, None)[0]


def test_rollout_gradient_matches_central_differences():
    d = _small(120, seed=62)
    dt, steps = 2e-13, 12
    xb, vb, qb, _ = oracle.rollout_grad_f64(d, dt, steps, lambda x, v: _divergence(x, v)[1:])
    rng = np.random.default_rng(3)
    for k in rng.choice(120, 4, replace=False):
        for key, grad, h in (("q", qb[k], 1e-6), ("vx", vb[k, 0], 1e-3), ("vz", vb[k, 2], 1e-3)):
            dp, dm = dict(d), dict(d)
            dp[key] = d[key].copy()
            dm[key] = d[key].copy()
            dp[key][k] += h
            dm[key][k] -= h
            fd = (_objective(dp, dt, steps) - _objective(dm, dt, steps)) / (2 * h)
            scale = np.max(np.abs(qb)) if key == "q" else np.max(np.abs(vb))
            assert abs(grad - fd) <= 1e-5 * abs(fd) + 1e-7 * scale, (key, k, grad, fd)
