"""Batched plumes through the C ABI: one persistent launch over instances of
uneven N equals the per-instance direct sums (FP64 to 1e-12; FP32C to the
compensated-FP32 bar against FP64), the autotuner reports every candidate
and keeps the fastest, and graph-captured leapfrog steps reproduce the
uncaptured ones bit for bit and the per-instance leapfrog runs to 1e-10."""
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


def _batch(sizes, seed=81):
    parts = []
    for k, n in enumerate(sizes):
        d = synth.cone(n, seed=seed + k) if k % 2 == 0 else synth.hemisphere(n, seed=seed + k)
        d["status"][::29] = 2
        d["status"][::53] = 0
        parts.append(d)
    keys = ("x", "y", "z", "vx", "vy", "vz", "q", "status", "species", "t_inj")
    allb = {k: np.concatenate([p[k] for p in parts]) for k in keys}
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    return parts, allb, off


def _per_instance(eng, parts, precision):
    import solveit_hip
    import torch
    out = []
    for d in parts:
        st = solveit_hip.DeviceState.from_numpy(d)
        torch.cuda.synchronize()
        out.append(torch.stack(eng.direct_dev(st, precision=precision), 1).cpu().numpy())
    return np.concatenate(out)


def test_batch_forces_match_per_instance(eng):
    import solveit_hip
    import torch
    sizes = [700, 3100, 1500, 257, 4096, 1]
    parts, allb, off = _batch(sizes)
    st = solveit_hip.DeviceState.from_numpy(allb)
    torch.cuda.synchronize()
    print("case batch forces", flush=True)
    ref64 = _per_instance(eng, parts, solveit_hip.SHP_PREC_FP64)
    plan = eng.batch(off, solveit_hip.SHP_PREC_FP64)
    for cfg in range(plan.N_CONFIGS):
        plan.set_config(cfg)
        got = torch.stack(plan.forces(st), 1).cpu().numpy()
        assert np.linalg.norm(got - ref64) <= 1e-12 * np.linalg.norm(ref64), cfg
    plan.close()
    plan = eng.batch(off, solveit_hip.SHP_PREC_FP32C)
    ms, chosen = plan.tune(st)
    print(f"  tuner ms {['%.3f' % m for m in ms]} chosen {chosen}")
    assert len(ms) == plan.N_CONFIGS and chosen == int(np.argmin(ms))
    got = torch.stack(plan.forces(st), 1).cpu().numpy()
    mask = allb["status"] == 1
    l2, mx, _ = oracle.rel_l2(got, ref64, mask)
    ref32 = _per_instance(eng, parts, solveit_hip.SHP_PREC_FP32C)
    l2i = oracle.rel_l2(ref32, ref64, mask)[0]
    print(f"  FP32C batch vs FP64: rel L2 {l2:.3e} max {mx:.3e}; per-instance FP32C {l2i:.3e}")
    assert l2 < 5e-6
    assert np.all(got[~mask] == 0)
    plan.close()


def test_batch_leapfrog_graph_and_per_instance(eng):
    import solveit_hip
    import torch
    sizes = [900, 2000, 333, 1600]
    parts, allb, off = _batch(sizes, seed=91)
    dt, steps = 1e-12, 12
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64)
    print("case batch leapfrog", flush=True)
    plan = eng.batch(off, solveit_hip.SHP_PREC_FP64)
    res = {}
    for g in (False, True):
        st = solveit_hip.DeviceState.from_numpy(allb)
        torch.cuda.synchronize()
        plan.leapfrog(st, opts, 0.0, dt, steps, use_graph=g)
        res[g] = st.to_numpy()
    for k in ("x", "y", "z", "vx", "vy", "vz", "status"):
        np.testing.assert_array_equal(res[True][k], res[False][k])
    ref = []
    for d in parts:
        st = solveit_hip.DeviceState.from_numpy(d)
        torch.cuda.synchronize()
        eng.leapfrog_run(st, opts, 0.0, dt, steps)
        eng.sync()
        ref.append(st.to_numpy())
    for k in ("x", "y", "z", "vx", "vy", "vz"):
        r = np.concatenate([q[k] for q in ref])
        scale = np.max(np.abs(r))
        assert np.max(np.abs(res[True][k] - r)) <= 1e-10 * scale, k
    plan.close()


def _batch_rollout_fn():
    import solveit_hip
    import torch

    def make_state(base, q, vx, vy, vz):
        return solveit_hip.DeviceState(base["x"].clone(), base["y"].clone(), base["z"].clone(),
                                       q.detach().clone(), base["status"], base["species"],
                                       vx.detach().clone(), vy.detach().clone(),
                                       vz.detach().clone(), torch.zeros_like(q))

    def copy(st):
        return solveit_hip.DeviceState(st.x.clone(), st.y.clone(), st.z.clone(), st.q, st.status,
                                       st.species, st.vx.clone(), st.vy.clone(), st.vz.clone(),
                                       st.t_inj)

    class BatchRollout(torch.autograd.Function):
        """Batched leapfrog rollout (graph-captured steps) with a checkpointed
        discrete adjoint through shp_batch_step_vjp."""

        @staticmethod
        def forward(ctx, q, vx, vy, vz, plan, base, opts, dt, nsteps, K):
            st = make_state(base, q, vx, vy, vz)
            ckpt = []
            for s0 in range(0, nsteps, K):
                ckpt.append((s0, copy(st)))
                plan.leapfrog(st, opts, s0 * dt, dt, min(K, nsteps - s0), use_graph=True)
            ctx.plan, ctx.opts, ctx.dt, ctx.nsteps, ctx.K, ctx.ckpt = plan, opts, dt, nsteps, K, ckpt
            return st.vx.clone(), st.vy.clone(), st.vz.clone()

        @staticmethod
        def backward(ctx, gvx, gvy, gvz):
            plan, dt = ctx.plan, ctx.dt
            xb = torch.zeros(3, gvx.shape[0], dtype=torch.float64, device=gvx.device)
            vb = torch.stack([gvx, gvy, gvz]).contiguous()
            qb = torch.zeros_like(gvx)
            for s0, cp in reversed(ctx.ckpt):
                st = copy(cp)
                starts = []
                for k in range(min(ctx.K, ctx.nsteps - s0)):
                    starts.append(copy(st))
                    plan.leapfrog(st, ctx.opts, (s0 + k) * dt, dt, 1, use_graph=False)
                for s in reversed(starts):
                    xb, vb = plan.step_vjp(s, dt, xb, vb, qb)
            return qb, vb[0], vb[1], vb[2], None, None, None, None, None, None

    return BatchRollout


def _sin2(vx, vy, vz, mask):
    s = (vx * vx + vy * vy) / (vx * vx + vy * vy + vz * vz)
    return s[mask].mean()


@pytest.mark.parametrize("prec", ["fp64", "fp32c"])
def test_batched_adjoint_directional_fd(eng, prec):
    """Gradient of the mean sin^2 of the velocity divergence angle with respect
    to charges and initial velocities over a batch of uneven plumes (K steps
    per checkpoint), against central differences of the FP64 rollout on
    seeded directions."""
    import solveit_hip
    import torch
    sizes = [600, 1500, 250, 1100, 900]
    parts, allb, off = _batch(sizes, seed=101)
    allb["status"][allb["status"] == 0] = 1
    base = {k: _t(allb[k]) for k in ("x", "y", "z")}
    base["status"] = _t(allb["status"], torch.int32)
    base["species"] = _t(allb["species"], torch.int32)
    dt, nsteps, K = 1e-13, 30, 10
    p = solveit_hip.SHP_PREC_FP64 if prec == "fp64" else solveit_hip.SHP_PREC_FP32C
    plan = eng.batch(off, p)
    plan64 = eng.batch(off, solveit_hip.SHP_PREC_FP64)
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT, precision=p)
    opts64 = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                        precision=solveit_hip.SHP_PREC_FP64)
    mask = base["status"] == 1
    Roll = _batch_rollout_fn()
    q = _t(allb["q"]).requires_grad_(True)
    v = [_t(allb[k]).requires_grad_(True) for k in ("vx", "vy", "vz")]
    print(f"case batched adjoint {prec}", flush=True)
    J = _sin2(*Roll.apply(q, *v, plan, base, opts, dt, nsteps, K), mask)
    J.backward()
    grads = [q.grad] + [t.grad for t in v]

    def J64(qq, vv):
        with torch.no_grad():
            return float(_sin2(*Roll.apply(qq, *vv, plan64, base, opts64, dt, nsteps, K), mask))

    gen = torch.Generator(device="cpu").manual_seed(2026)
    h = 1e-6
    worst = 0.0
    for _ in range(8):
        dirs = [torch.randn(q.shape[0], generator=gen, dtype=torch.float64).to("cuda:0")
                for _ in range(4)]
        dq = dirs[0] * 0.01
        dv = [d * 10.0 for d in dirs[1:]]
        ana = float(sum((g * d).sum() for g, d in zip(grads, [dq] + dv)))
        jp = J64(q.detach() + h * dq, [t.detach() + h * d for t, d in zip(v, dv)])
        jm = J64(q.detach() - h * dq, [t.detach() - h * d for t, d in zip(v, dv)])
        fd = (jp - jm) / (2 * h)
        worst = max(worst, abs(ana - fd) / abs(fd))
        print(f"  directional: adjoint {ana:.8e} fd {fd:.8e}")
    print(f"  worst relative difference {worst:.2e}")
    assert worst < (1e-5 if prec == "fp64" else 1e-3)
    plan.close()
    plan64.close()


def test_batch_forces_and_leapfrog_match_oracle(eng):
    """Each instance of a batch against the oracle itself: the batched FP64
    forces equal oracle.direct_f64 to 1e-12 (FP32C within 5e-6), and 12
    graph-captured batched leapfrog steps equal oracle.leapfrog_run_f64 to
    1e-10 per instance."""
    import solveit_hip
    import torch
    sizes = [700, 1900, 257, 1300, 1]
    parts, allb, off = _batch(sizes, seed=111)
    st = solveit_hip.DeviceState.from_numpy(allb)
    torch.cuda.synchronize()
    print("case batch vs oracle forces", flush=True)
    ref = [oracle.direct_f64(d) for d in parts]
    for prec, tol in ((solveit_hip.SHP_PREC_FP64, 1e-12), (solveit_hip.SHP_PREC_FP32C, 5e-6)):
        plan = eng.batch(off, prec)
        got = torch.stack(plan.forces(st), 1).cpu().numpy()
        plan.close()
        for k, d in enumerate(parts):
            g = got[off[k]:off[k + 1]]
            m = d["status"] == 1
            if prec == solveit_hip.SHP_PREC_FP64:
                assert np.linalg.norm(g - ref[k]) <= tol * np.linalg.norm(ref[k]) + 1e-300, k
            elif m.sum() > 1:
                assert oracle.rel_l2(g, ref[k], m)[0] <= tol, k
    print("case batch vs oracle leapfrog", flush=True)
    dt, steps = 1e-12, 12
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64)
    plan = eng.batch(off, solveit_hip.SHP_PREC_FP64)
    st = solveit_hip.DeviceState.from_numpy(allb)
    torch.cuda.synchronize()
    plan.leapfrog(st, opts, 0.0, dt, steps, use_graph=True)
    got = st.to_numpy()
    plan.close()
    for k, d in enumerate(parts):
        r, _ = oracle.leapfrog_run_f64(d, 0.0, dt, steps)
        for c in ("x", "y", "z", "vx", "vy", "vz"):
            g = got[c][off[k]:off[k + 1]]
            scale = np.max(np.abs(r[c])) + 1e-300
            assert np.max(np.abs(g - r[c])) <= 1e-10 * scale, (k, c)
        np.testing.assert_array_equal(got["status"][off[k]:off[k + 1]], r["status"])


def test_batched_adjoint_matches_oracle_at_the_200_step_schedule(eng):
    """The batched checkpointed adjoint at the 64-plume workload's schedule
    (K = 200 leapfrog steps, a checkpoint every 20) on 8 plumes of uneven N drawn
    lognormally in [256, 2048]: the gradient of the mean sin^2 of the
    velocity divergence angle with respect to charges and initial velocities
    equals, instance by instance, the oracle's FP64 discrete adjoint
    (oracle.rollout_grad_f64) to 1e-8."""
    from concurrent.futures import ThreadPoolExecutor
    import solveit_hip
    import torch
    rng = np.random.default_rng(5)
    sizes = [int(v) for v in np.clip(np.round(np.exp(rng.normal(np.log(700), 0.6, 8))), 256, 2048)]
    parts, allb, off = _batch(sizes, seed=121)
    for d in parts:
        d["status"][d["status"] == 0] = 1
    allb["status"][allb["status"] == 0] = 1
    base = {k: _t(allb[k]) for k in ("x", "y", "z")}
    base["status"] = _t(allb["status"], torch.int32)
    base["species"] = _t(allb["species"], torch.int32)
    dt, nsteps, K = 1e-14, 200, 20
    plan = eng.batch(off, solveit_hip.SHP_PREC_FP64)
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64)
    mask = base["status"] == 1
    Roll = _batch_rollout_fn()
    q = _t(allb["q"]).requires_grad_(True)
    v = [_t(allb[k]).requires_grad_(True) for k in ("vx", "vy", "vz")]
    print(f"case batched adjoint vs oracle, sizes {sizes}", flush=True)
    J = _sin2(*Roll.apply(q, *v, plan, base, opts, dt, nsteps, K), mask)
    J.backward()
    plan.close()
    gq = q.grad.cpu().numpy()
    gv = np.stack([t.grad.cpu().numpy() for t in v], 1)
    M = int(mask.sum().item())

    def oracle_grad(k):
        d = parts[k]
        kick = d["status"] == 1

        def gfin(x, vel):
            vx, vy, vz = vel[:, 0], vel[:, 1], vel[:, 2]
            a = vx * vx + vy * vy
            b = a + vz * vz
            g = np.stack([2 * vx * vz * vz / b ** 2, 2 * vy * vz * vz / b ** 2, -2 * a * vz / b ** 2], 1)
            return np.zeros_like(x), np.where(kick[:, None], g / M, 0.0)

        return oracle.rollout_grad_f64(d, dt, nsteps, gfin)

    with ThreadPoolExecutor(8) as ex:
        res = list(ex.map(oracle_grad, range(len(parts))))
    for k, (xb, vb, qb, _) in enumerate(res):
        sl = slice(off[k], off[k + 1])
        assert np.linalg.norm(gq[sl] - qb) <= 1e-8 * np.linalg.norm(qb), k
        assert np.linalg.norm(gv[sl] - vb) <= 1e-8 * np.linalg.norm(vb), k
    print(f"  {len(parts)} instances, {sum(sizes)} ions: adjoint equals the oracle's")
