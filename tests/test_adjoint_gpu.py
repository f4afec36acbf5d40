"""GPU force JVP/VJP and leapfrog adjoint through the C ABI.

The JVP and VJP match the oracle's FP64 derivatives (relative 1e-10).  A
torch.autograd.Function (harness only) runs the leapfrog rollout forward
through shp_leapfrog_run, checkpoints the state every K steps and runs the
discrete adjoint backward with shp_leapfrog_step_vjp_dev, recomputing each
segment from its checkpoint; the gradient of a plume-divergence objective
with respect to per-ion charge and initial (injection) velocity matches the
oracle's FP64 adjoint and central finite differences of the GPU rollout."""
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


def _t(a, dtype=None):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype or torch.float64, device="cuda:0")


def _case(n=600, seed=71):
    d = synth.hemisphere(n, seed=seed)
    d["status"][::11] = 2
    d["q"][::7] = -1.0
    d["species"][::7] = 5
    return d


def test_force_jvp_vjp_match_oracle(eng):
    import solveit_hip
    import torch
    d = _case()
    n = d["x"].shape[0]
    rng = np.random.default_rng(5)
    dx, dy, dz = (rng.normal(0, 1e-9, n) for _ in range(3))
    dq = rng.normal(0, 0.1, n)
    b = rng.normal(0, 1, (n, 3))
    print("case force jvp/vjp", flush=True)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    jv = torch.stack(eng.force_jvp(st, _t(dx), _t(dy), _t(dz), _t(dq)), 1).cpu().numpy()
    ref = oracle.force_jvp_f64(d, dx, dy, dz, dq)
    assert np.linalg.norm(jv - ref) <= 1e-10 * np.linalg.norm(ref)
    (xb, yb, zb), qb = eng.force_vjp(st, _t(b[:, 0]), _t(b[:, 1]), _t(b[:, 2]))
    rxb, rqb = oracle.force_vjp_f64(d, b)
    got = np.stack([xb.cpu().numpy(), yb.cpu().numpy(), zb.cpu().numpy()], 1)
    assert np.linalg.norm(got - rxb) <= 1e-10 * np.linalg.norm(rxb)
    assert np.linalg.norm(qb.cpu().numpy() - rqb) <= 1e-10 * np.linalg.norm(rqb)


def _make_state(base, q, vx, vy, vz):
    import solveit_hip
    import torch
    z = torch.zeros_like(q)
    return solveit_hip.DeviceState(base["x"].clone(), base["y"].clone(), base["z"].clone(),
                                   q.detach().clone(), base["status"], base["species"],
                                   vx.detach().clone(), vy.detach().clone(), vz.detach().clone(),
                                   z)


def _copy_state(st):
    import solveit_hip
    return solveit_hip.DeviceState(st.x.clone(), st.y.clone(), st.z.clone(), st.q, st.status,
                                   st.species, st.vx.clone(), st.vy.clone(), st.vz.clone(),
                                   st.t_inj)


def _rollout_fn():
    import solveit_hip
    import torch

    class PlumeRollout(torch.autograd.Function):
        """Leapfrog rollout (FP64 direct force) with a checkpointed discrete adjoint."""

        @staticmethod
        def forward(ctx, q, vx, vy, vz, eng, base, dt, nsteps, K):
            opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                              precision=solveit_hip.SHP_PREC_FP64)
            st = _make_state(base, q, vx, vy, vz)
            ckpt = []
            for s0 in range(0, nsteps, K):
                ckpt.append((s0, _copy_state(st)))
                eng.leapfrog_run(st, opts, s0 * dt, dt, min(K, nsteps - s0))
            eng.sync()
            ctx.eng, ctx.opts, ctx.dt, ctx.nsteps, ctx.ckpt = eng, opts, dt, nsteps, ckpt
            ctx.K = K
            return st.x.clone(), st.y.clone(), st.z.clone()

        @staticmethod
        def backward(ctx, gx, gy, gz):
            eng, dt = ctx.eng, ctx.dt
            xb = (gx.contiguous(), gy.contiguous(), gz.contiguous())
            vb = tuple(torch.zeros_like(gx) for _ in range(3))
            qb = torch.zeros_like(gx)
            for s0, cp in reversed(ctx.ckpt):
                steps = min(ctx.K, ctx.nsteps - s0)
                st = _copy_state(cp)
                starts = []
                for k in range(steps):  # recompute the segment from its checkpoint
                    starts.append(_copy_state(st))
                    eng.leapfrog_run(st, ctx.opts, (s0 + k) * dt, dt, 1)
                for s in reversed(starts):
                    xb, vb = eng.leapfrog_step_vjp(s, dt, xb, vb, qb)
            return qb, vb[0], vb[1], vb[2], None, None, None, None, None

    return PlumeRollout


def _divergence(x, y):
    return (x * x + y * y).mean()


def test_rollout_gradient_matches_oracle_and_fd(eng):
    import torch
    d = _case(400, seed=72)
    dt, nsteps, K = 2e-13, 24, 5
    base = {k: _t(d[k]) for k in ("x", "y", "z")}
    base["status"] = _t(d["status"], torch.int32)
    base["species"] = _t(d["species"], torch.int32)
    q = _t(d["q"]).requires_grad_(True)
    vx, vy, vz = (_t(d[k]).requires_grad_(True) for k in ("vx", "vy", "vz"))
    Roll = _rollout_fn()
    print("case rollout adjoint", flush=True)
    x1, y1, z1 = Roll.apply(q, vx, vy, vz, eng, base, dt, nsteps, K)
    J = _divergence(x1, y1)
    J.backward()
    gq, gvx, gvz = q.grad.cpu().numpy(), vx.grad.cpu().numpy(), vz.grad.cpu().numpy()
    n = d["x"].shape[0]

    def gfin(x, v):
        g = np.zeros_like(x)
        g[:, 0] = 2 * x[:, 0] / n
        g[:, 1] = 2 * x[:, 1] / n
        return g, np.zeros_like(v)

    rxb, rvb, rqb, fin = oracle.rollout_grad_f64(d, dt, nsteps, gfin)
    np.testing.assert_allclose(x1.detach().cpu().numpy(), fin["x"], rtol=0, atol=1e-10 * np.max(np.abs(fin["x"])))
    assert np.linalg.norm(gq - rqb) <= 1e-8 * np.linalg.norm(rqb)
    assert np.linalg.norm(gvx - rvb[:, 0]) <= 1e-8 * np.linalg.norm(rvb[:, 0])
    assert np.linalg.norm(gvz - rvb[:, 2]) <= 1e-8 * np.linalg.norm(rvb[:, 2])

    def objective(qq, vvx):
        with torch.no_grad():
            a, b, _ = Roll.apply(qq, vvx, vy.detach(), vz.detach(), eng, base, dt, nsteps, K)
            return float(_divergence(a, b))

    rng = np.random.default_rng(6)
    for k in rng.choice(n, 3, replace=False):
        for which, h in (("q", 1e-6), ("vx", 1e-3)):
            qp, qm = q.detach().clone(), q.detach().clone()
            vp, vm = vx.detach().clone(), vx.detach().clone()
            if which == "q":
                qp[k] += h
                qm[k] -= h
                g = gq[k]
                scale = np.max(np.abs(gq))
            else:
                vp[k] += h
                vm[k] -= h
                g = gvx[k]
                scale = np.max(np.abs(gvx))
            fd = (objective(qp, vp) - objective(qm, vm)) / (2 * h)
            print(f"  {which}[{k}]: adjoint {g:.6e} fd {fd:.6e}")
            assert abs(g - fd) <= 1e-5 * abs(fd) + 1e-7 * scale
