"""Energies on the GPU against the oracle (FP64, relative 1e-12), and total
energy drift of leapfrog runs: FP64 direct forces conserve KE + PE to the
integrator's O(dt^2) level; FMM forces (FP32 expansions and pairs) are
reported and bounded."""
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


def test_energy_matches_oracle(eng):
    import solveit_hip
    import torch
    d = synth.cone(4096, seed=51)
    d["status"][::19] = 2
    d["status"][::43] = 0
    print("case energy n=4096", flush=True)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    ke, pe = eng.energy(st)
    rke, rpe = oracle.energy_f64(d)
    print(f"  KE {ke:.6e} ({ke / rke - 1:.2e}), PE {pe:.6e} ({pe / rpe - 1:.2e})")
    assert abs(ke - rke) <= 1e-12 * abs(rke)
    assert abs(pe - rpe) <= 1e-12 * abs(rpe)


def test_field_energy_matches_oracle(eng):
    import torch
    mesh = synth.tip_field_mesh()
    eng.efield_setup(mesh)
    d0 = synth.cone(2000, seed=52)
    rng = np.random.default_rng(53)
    p0 = (d0["x"], d0["y"], d0["z"])
    p1 = (d0["x"] + rng.normal(0, 2e-8, 2000), d0["y"] + rng.normal(0, 2e-8, 2000),
          d0["z"] + rng.uniform(0, 2e-7, 2000))
    ns = rng.integers(2, 9, 2000).astype(np.int32)
    print("case field energy n=2000", flush=True)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda:0")
    fe = eng.field_energy(*(t(a) for a in p0), *(t(a) for a in p1), t(d0["q"]),
                          nsteps=torch.as_tensor(ns, device="cuda:0"))
    got = fe.cpu().numpy()
    ref = oracle.field_energy_f64(mesh, p0, p1, d0["q"], nsteps=ns)
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(got - ref)) <= 1e-12 * scale


def _run_energy(eng, d, method, steps, dt, **kw):
    import solveit_hip
    import torch
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    e0 = sum(eng.energy(st))
    opts = solveit_hip.make_step_opts(method=method, **kw)
    eng.leapfrog_run(st, opts, 0.0, dt, steps)
    eng.sync()
    e1 = sum(eng.energy(st))
    return e0, e1


def test_energy_drift_direct_and_fmm(eng):
    """Dense 20 nm hemisphere (Coulomb explosion), same physical time with dt
    and dt/2: the FP64-direct drift of KE + PE falls about fourfold (a
    second-order integrator); FMM forces drift by about as much."""
    import solveit_hip
    d = synth.hemisphere(3000, seed=54)
    print("case energy drift", flush=True)
    drifts = []
    for steps, dt in ((50, 4e-14), (100, 2e-14)):
        e0, e1 = _run_energy(eng, d, solveit_hip.SHP_METHOD_DIRECT, steps, dt,
                             precision=solveit_hip.SHP_PREC_FP64)
        drifts.append((e1 - e0) / abs(e0))
    f0, f1 = _run_energy(eng, d, solveit_hip.SHP_METHOD_FMM, 100, 2e-14, fmm_order=10,
                         leaf_size=20)
    drift_fmm = (f1 - f0) / abs(f0)
    print(f"  FP64 direct drift dt=4e-14 {drifts[0]:.3e}, dt=2e-14 {drifts[1]:.3e} "
          f"(ratio {drifts[0] / drifts[1]:.2f}); FMM dt=2e-14 {drift_fmm:.3e}")
    assert 2.5 < drifts[0] / drifts[1] < 6.0
    assert abs(drifts[1]) < 1e-3
    assert abs(drift_fmm - drifts[1]) < 0.1 * abs(drifts[1]) + 1e-6
