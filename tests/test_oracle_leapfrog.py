"""Pins the oracle's DKD leapfrog (CPU only): exact parabola under a constant
field, time reversibility, injection clamps, boundary conditions, folding."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _uniform_mesh(ez=1.0):
    m = synth.tip_field_mesh(nr=8, nz=8)
    m["er"] = np.zeros_like(m["er"])
    m["ez"] = np.full_like(m["ez"], ez)
    return m


def _one_ion(z0=1e-7, vz=500.0, t_inj=0.0, q=1.0, species=0):
    return {"x": np.array([2e-8]), "y": np.array([-1e-8]), "z": np.array([z0]),
            "vx": np.array([10.0]), "vy": np.array([-5.0]), "vz": np.array([vz]),
            "q": np.array([q]), "status": np.array([1], np.int32),
            "species": np.array([species], np.int32), "t_inj": np.array([t_inj])}


def test_constant_field_gives_exact_parabola():
    m = _uniform_mesh()
    d = _one_ion()
    dt, k = 1e-12, 50
    s, _ = oracle.leapfrog_run_f64(d, 0.0, dt, k, mesh=m)
    a = oracle.lib.orc_e_over_amu() * 1.0 / 111.0 * m["E0"]
    T = dt * k
    np.testing.assert_allclose(s["z"][0], d["z"][0] + d["vz"][0] * T + 0.5 * a * T * T, rtol=1e-13)
    np.testing.assert_allclose(s["vz"][0], d["vz"][0] + a * T, rtol=1e-13)
    np.testing.assert_allclose(s["x"][0], d["x"][0] + d["vx"][0] * T, rtol=1e-13)


def test_injection_clamp_hand_case():
    # injected at 0.7 dt: the first half drift (to 0.5 dt) leaves it alone, the
    # kick is skipped (it is outside the range at 0.5 dt), and the second half
    # drift moves it by (dt - 0.7 dt) v.
    m = _uniform_mesh()
    dt = 1e-12
    d = _one_ion(t_inj=0.7 * dt)
    s, _ = oracle.leapfrog_run_f64(d, 0.0, dt, 1, mesh=m)
    np.testing.assert_allclose(s["z"][0], d["z"][0] + d["vz"][0] * 0.3 * dt, rtol=1e-14)
    assert s["vz"][0] == d["vz"][0]
    # injected at 0.2 dt: drift (0.3 dt), kick over 0.8 dt, drift 0.5 dt
    d = _one_ion(t_inj=0.2 * dt)
    s, _ = oracle.leapfrog_run_f64(d, 0.0, dt, 1, mesh=m)
    a = oracle.lib.orc_e_over_amu() / 111.0 * m["E0"]
    v1 = d["vz"][0] + a * 0.8 * dt
    np.testing.assert_allclose(s["vz"][0], v1, rtol=1e-14)
    np.testing.assert_allclose(s["z"][0], d["z"][0] + d["vz"][0] * 0.3 * dt + v1 * 0.5 * dt,
                               rtol=1e-14)


def test_time_reversibility_coulomb():
    d = synth.hemisphere(64, seed=9)
    dt, k = 1e-12, 40
    s, _ = oracle.leapfrog_run_f64(d, 0.0, dt, k)
    back = dict(s)
    for c in ("vx", "vy", "vz"):
        back[c] = -s[c]
    r, _ = oracle.leapfrog_run_f64(back, 0.0, dt, k)
    pos0 = np.stack([d["x"], d["y"], d["z"]], 1)
    pos1 = np.stack([r["x"], r["y"], r["z"]], 1)
    assert np.max(np.abs(pos1 - pos0)) < 1e-12 * np.max(np.abs(pos0))
    np.testing.assert_allclose(-r["vz"], d["vz"], rtol=1e-9, atol=1e-9)


def test_energy_is_bounded_for_coulomb_cluster():
    d = synth.hemisphere(128, seed=4)
    K = oracle.k_const()

    def energy(s):
        m = np.array([oracle.species_mass(x) for x in s["species"]])
        ke = 0.5 * np.sum(m * (s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2))
        p = np.stack([s["x"], s["y"], s["z"]], 1)
        r2 = np.sum((p[:, None, :] - p[None, :, :]) ** 2, -1) + 6.4e-19
        qq = s["q"][:, None] * s["q"][None, :]
        np.fill_diagonal(qq, 0.0)
        pe = 0.5 * K * np.sum(qq / np.sqrt(r2))
        return ke + pe   # in amu m^2/s^2 (K carries 1/amu)

    e0 = energy(d)
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-13, 100)
    assert abs(energy(s) - e0) < 1e-3 * abs(e0)


def test_spatial_and_temporal_boundaries():
    d = _one_ion(z0=1e-6, vz=1e6)
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-12, 3, bc_space=(0, 0, 1.002e-6, 0))
    assert s["status"][0] == 0
    d = _one_ion()
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-12, 3, bc_time=2.2e-12)
    assert s["status"][0] == 0
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-12, 2, bc_time=2.2e-12)
    assert s["status"][0] == 1


@pytest.mark.parametrize("x,y", [(-1.0, 2.0), (-1.0, -2.0), (1.0, -2.0), (3.0, 4.0)])
def test_four_fold_folding_matches_reference_rules(x, y):
    d = _one_ion(q=0.0)  # uncharged: its own sector images exert no force
    d["x"][:] = x * 1e-8
    d["y"][:] = y * 1e-8
    d["vx"][:] = 0.0
    d["vy"][:] = 0.0
    d["vz"][:] = 0.0
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-12, 1, sym="rotational", nsec=4)
    # rotateToFirstQuadrant (particles.cpp:232-262)
    if x < 0 and y < 0:
        ex, ey = -x, -y
    elif x < 0:
        ex, ey = y, -x
    elif y < 0:
        ex, ey = -y, x
    else:
        ex, ey = x, y
    np.testing.assert_allclose([s["x"][0], s["y"][0]], [ex * 1e-8, ey * 1e-8], rtol=1e-15)
    assert s["x"][0] >= 0 and s["y"][0] >= 0


def test_translational_wrap():
    d = _one_ion()
    d["x"][:] = 1.3e-6
    d["vx"][:] = d["vy"][:] = d["vz"][:] = 0.0
    s, _ = oracle.leapfrog_run_f64(d, 0.0, 1e-12, 1, sym="translational",
                                   lattice=np.zeros((1, 3), np.int32), domain=(2e-6, 0, 0))
    np.testing.assert_allclose(s["x"][0], 1.3e-6 - 2e-6, rtol=1e-12)
