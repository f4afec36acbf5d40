"""Parity of the HIP background field and DKD leapfrog with the oracle (FP64),
through the C ABI.  Bar: 1e-12 relative for the field; trajectories of a few
tens of steps within 1e-10 relative (double sums in a different order, carried
through the integration)."""
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


@pytest.fixture(scope="module")
def mesh():
    return synth.tip_field_mesh()


def _t(a, dtype="f64"):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64 if dtype == "f64"
                           else torch.int32, device="cuda:0")


def test_efield_matches_oracle(eng, mesh):
    import torch
    eng.efield_setup(mesh)
    rng = np.random.default_rng(5)
    d = synth.cone(20000, seed=6)
    # plume points plus points spread over the whole hull and some gated ones
    n2 = 5000
    rr = 140 * mesh["r0"] * rng.random(n2) ** 2
    ph = rng.uniform(0, 2 * np.pi, n2)
    zz = 140 * mesh["r0"] * rng.random(n2) ** 2
    x = np.concatenate([d["x"], rr * np.cos(ph), [0.0, 1e-7, 0.0, 3e-6]])
    y = np.concatenate([d["y"], rr * np.sin(ph), [0.0, 0.0, 0.0, 0.0]])
    z = np.concatenate([d["z"], zz, [0.0, -1e-9, 141 * mesh["r0"], 0.0]])
    print("case efield n=%d" % x.shape[0], flush=True)
    (ex, ey, ez), tri = eng.efield_dev(_t(x), _t(y), _t(z), want_tri=True)
    eng.sync()
    got = torch.stack([ex, ey, ez], 1).cpu().numpy()
    ref, edge, _ = oracle.efield_f64(mesh, x, y, z)
    assert np.all(edge == 0)
    scale = np.maximum(np.linalg.norm(ref, axis=1), 1e-300)
    err = np.linalg.norm(got - ref, axis=1) / scale
    assert np.all(np.linalg.norm(ref, axis=1)[-3:-1] == 0) and np.all(got[-3:-1] == 0)
    assert err.max() <= 1e-12


def _run_both(eng, mesh, d, nsteps, dt, **kw):
    import solveit_hip
    import torch
    symkind = kw.pop("sym", "none")
    nsec = kw.pop("nsec", 4)
    lattice = kw.pop("lattice", None)
    domain = kw.pop("domain", (0.0, 0.0, 0.0))
    bc_space = kw.pop("bc_space", (0, 0, 0, 0))
    bc_time = kw.pop("bc_time", 0.0)
    use_field = mesh is not None
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    if use_field:
        eng.efield_setup(mesh)
    sym = solveit_hip.make_symmetry(symkind, nsec=nsec, lattice=lattice, domain=domain)
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64, symmetry=sym,
                                      use_efield=use_field, bc_space=bc_space, bc_time=bc_time)
    acc = eng.leapfrog_run(st, opts, 0.0, dt, nsteps)
    eng.sync()
    got = st.to_numpy()
    ref, racc = oracle.leapfrog_run_f64(d, 0.0, dt, nsteps, mesh=mesh, sym=symkind, nsec=nsec,
                                        lattice=lattice, domain=domain, bc_space=bc_space,
                                        bc_time=bc_time)
    return got, ref, torch.stack(acc, 1).cpu().numpy(), racc


def _close(got, ref, keys=("x", "y", "z", "vx", "vy", "vz"), tol=1e-10):
    for k in keys:
        scale = np.max(np.abs(ref[k])) + 1e-300
        assert np.max(np.abs(got[k] - ref[k])) <= tol * scale, k


def test_leapfrog_with_field_matches_oracle(eng, mesh):
    d = synth.cone(1500, seed=8)
    print("case leapfrog field", flush=True)
    got, ref, acc, racc = _run_both(eng, mesh, d, 20, 1e-12)
    _close(got, ref)
    np.testing.assert_array_equal(got["status"], ref["status"])


def test_leapfrog_injection_and_boundaries(eng, mesh):
    d = synth.cone(1200, seed=9)
    dt = 1e-12
    d["t_inj"] = np.sort(np.random.default_rng(1).uniform(0, 10 * dt, 1200))
    print("case leapfrog injection/bc", flush=True)
    got, ref, _, _ = _run_both(eng, mesh, d, 15, dt, bc_space=(0, 0, 9.0e-6, 2e-7),
                               bc_time=12 * dt)
    np.testing.assert_array_equal(got["status"], ref["status"])
    assert np.any(ref["status"] == 0) and np.any(ref["status"] == 1)
    _close(got, ref)


@pytest.mark.parametrize("nsec", [4, 6])
def test_leapfrog_rotational(eng, mesh, nsec):
    d = synth.cone(800, seed=10)
    print(f"case leapfrog rot {nsec}", flush=True)
    got, ref, _, _ = _run_both(eng, None, d, 10, 1e-12, sym="rotational", nsec=nsec)
    _close(got, ref)


def test_leapfrog_translational(eng):
    d = synth.cell_fill(600, seed=11)
    lat = synth.full_lattice(1, (1, 1, 1))
    print("case leapfrog trans", flush=True)
    got, ref, _, _ = _run_both(eng, None, d, 10, 1e-12, sym="translational", lattice=lat,
                               domain=(2e-6, 2e-6, 2e-6))
    _close(got, ref)


def test_leapfrog_fp32c_trajectory(eng, mesh):
    """FP32C forces with FP64 state against FP64 forces over 50 steps: the
    typical ion's velocity agrees to float-force accuracy (close encounters
    amplify the per-pair float error, so the maximum is reported, not bounded
    tightly)."""
    import solveit_hip
    import torch
    d = synth.cone(4096, seed=12)
    out = {}
    for prec in (solveit_hip.SHP_PREC_FP64, solveit_hip.SHP_PREC_FP32C):
        st = solveit_hip.DeviceState.from_numpy(d)
        torch.cuda.synchronize()
        eng.efield_setup(mesh)
        opts = solveit_hip.make_step_opts(precision=prec, use_efield=True)
        eng.leapfrog_run(st, opts, 0.0, 1e-12, 50)
        eng.sync()
        out[prec] = st.to_numpy()
    a, b = out[solveit_hip.SHP_PREC_FP64], out[solveit_hip.SHP_PREC_FP32C]
    dv = np.sqrt((a["vx"] - b["vx"]) ** 2 + (a["vy"] - b["vy"]) ** 2 + (a["vz"] - b["vz"]) ** 2)
    v = np.sqrt(a["vx"] ** 2 + a["vy"] ** 2 + a["vz"] ** 2)
    r = dv / v
    print(f"  fp32c vs fp64 velocity rel diff after 50 steps: median {np.median(r):.3e} "
          f"p99 {np.percentile(r, 99):.3e} max {np.max(r):.3e}")
    assert np.median(r) < 1e-7
    assert np.percentile(r, 99) < 1e-4


def test_resident_state_on_host_arrays(eng, mesh):
    """shp_state_upload_host / shp_state_run / shp_state_download_host (the
    HipLeapfrogDKD adapter's path on the driver's float arrays) equal
    shp_leapfrog_run on the same FP64 state, rounded to float, with injection
    and boundary conditions."""
    import solveit_hip
    import torch
    d = synth.cone(1500, seed=12)
    dt = 1e-12
    d["t_inj"] = np.sort(np.random.default_rng(2).uniform(0, 6 * dt, 1500))
    for k in ("x", "y", "z", "vx", "vy", "vz", "q", "t_inj"):
        d[k] = d[k].astype(np.float32).astype(np.float64)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64, use_efield=True,
                                      bc_space=(0, 0, 9.0e-6, 2e-7), bc_time=8 * dt)
    print("case resident state", flush=True)
    eng.state_upload_host(pos4, vel4, acc4, d["t_inj"])
    for k in range(3):  # one step per call, as the adapter runs them
        eng.state_run(opts, k * dt, dt, 1)
    eng.state_download_host(pos4, vel4, acc4)
    st = solveit_hip.DeviceState.from_numpy(d)
    torch.cuda.synchronize()
    acc = eng.leapfrog_run(st, opts, 0.0, dt, 1)
    for k in range(1, 3):
        acc = eng.leapfrog_run(st, opts, k * dt, dt, 1, acc=acc)
    ref = st.to_numpy()
    a = torch.stack(acc, 1).cpu().numpy()
    np.testing.assert_array_equal(pos4[:, 0], ref["x"].astype(np.float32))
    np.testing.assert_array_equal(pos4[:, 2], ref["z"].astype(np.float32))
    np.testing.assert_array_equal(vel4[:, 1], ref["vy"].astype(np.float32))
    np.testing.assert_array_equal(acc4[:, 3], ref["status"].astype(np.float32))
    np.testing.assert_array_equal(acc4[:, :3], a.astype(np.float32))
    np.testing.assert_array_equal(pos4[:, 3], d["q"].astype(np.float32))
    np.testing.assert_array_equal(vel4[:, 3], d["species"].astype(np.float32))
    assert np.any(ref["status"] == 0)


@pytest.mark.parametrize("kind", ["space", "space_and_time"])
def test_leapfrog_substeps_match_oracle(eng, mesh, kind):
    """LeapfrogDKDSubsteps on the GPU against the oracle's restatement: three
    levels (dt, dt/3, dt/9) with |z| limits (and time-since-injection limits),
    fast ions that cross the limits between steps, staggered injection, the
    tip field and boundary conditions; FP64 direct force.  Positions and
    velocities to 1e-10, statuses and levels exactly."""
    import solveit_hip
    import torch
    d = synth.cone(900, seed=13)
    dt = 1e-12
    d["vz"] = d["vz"] * 1000.0  # cross the 1 um and 3 um limits within a few steps
    d["t_inj"] = np.sort(np.random.default_rng(4).uniform(0, 4 * dt, 900))
    if kind == "space":
        levels = [(1, (0, 0, 0, 0), 0.0), (1, (0, 0, 3e-6, 0), 0.0), (1, (0, 0, 1e-6, 0), 0.0)]
    else:
        levels = [(3, (0, 0, 0, 0), 0.0), (3, (0, 0, 3e-6, 0), 4e-12), (3, (0, 0, 1e-6, 0), 1.5e-12)]
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64, use_efield=True,
                                      bc_space=(0, 0, 1.2e-5, 0), bc_time=0.0)
    print(f"case substeps {kind}", flush=True)
    eng.efield_setup(mesh)
    st = solveit_hip.DeviceState.from_numpy(d)
    level = torch.full((900,), 2, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    acc = eng.leapfrog_substeps(st, opts, levels, level, 0.0, dt, 6)
    got = st.to_numpy()
    ref, rlev, racc = oracle.leapfrog_substeps_f64(d, 0.0, dt, 6, levels, mesh=mesh,
                                                   bc_space=(0, 0, 1.2e-5, 0))
    glev = level.cpu().numpy()
    print(f"  levels: {np.bincount(glev + 1, minlength=4)} (-1, 0, 1, 2); statuses "
          f"{np.bincount(got['status'] + 1, minlength=4)}")
    np.testing.assert_array_equal(glev, rlev)
    np.testing.assert_array_equal(got["status"], ref["status"])
    assert len(set(glev.tolist())) >= 2  # ions on more than one level
    _close(got, ref)
    a = torch.stack(acc, 1).cpu().numpy()
    assert np.linalg.norm(a - racc) <= 1e-10 * np.linalg.norm(racc)


def test_resident_substeps_on_host_arrays(eng, mesh):
    """shp_state_run_substeps (the HipLeapfrogDKDSubsteps adapter's path, one
    step per call with the levels kept in the handle) equals
    shp_leapfrog_substeps_run on the same FP64 state, rounded to float."""
    import solveit_hip
    import torch
    d = synth.cone(800, seed=14)
    dt = 1e-12
    d["vz"] = d["vz"] * 1000.0
    for k in ("x", "y", "z", "vx", "vy", "vz", "q", "t_inj"):
        d[k] = d[k].astype(np.float32).astype(np.float64)
    levels = [(1, (0, 0, 0, 0), 0.0), (1, (0, 0, 3e-6, 0), 0.0), (1, (0, 0, 1e-6, 0), 0.0)]
    opts = solveit_hip.make_step_opts(method=solveit_hip.SHP_METHOD_DIRECT,
                                      precision=solveit_hip.SHP_PREC_FP64, use_efield=True)
    print("case resident substeps", flush=True)
    eng.efield_setup(mesh)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    eng.state_upload_host(pos4, vel4, acc4, d["t_inj"])
    for k in range(3):
        eng.state_run_substeps(opts, levels, k * dt, dt, 1, reset=(k == 0))
    eng.state_download_host(pos4, vel4, acc4)
    st = solveit_hip.DeviceState.from_numpy(d)
    level = torch.full((800,), 2, dtype=torch.int32, device="cuda:0")
    torch.cuda.synchronize()
    acc = None
    for k in range(3):
        acc = eng.leapfrog_substeps(st, opts, levels, level, k * dt, dt, 1, acc=acc)
    ref = st.to_numpy()
    np.testing.assert_array_equal(pos4[:, 2], ref["z"].astype(np.float32))
    np.testing.assert_array_equal(vel4[:, 0], ref["vx"].astype(np.float32))
    np.testing.assert_array_equal(acc4[:, 3], ref["status"].astype(np.float32))
    np.testing.assert_array_equal(acc4[:, :3], torch.stack(acc, 1).cpu().numpy().astype(np.float32))
