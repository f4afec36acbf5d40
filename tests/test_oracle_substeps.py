"""Pins the oracle's LeapfrogDKDSubsteps restatement (CPU only): one level is
LeapfrogDKD bit for bit; a lone ion in a uniform field follows the exact
parabola at every level; the reference's kick window (tempVelTime + dt_level)
on an ion injected mid-step, derived by hand; and the level rules of
checkConditions on hand cases (spatial, temporal, both, none admitted)."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth

NONE = (0, (0, 0, 0, 0), 0.0)


def _uniform_mesh(ez=1.0):
    m = synth.tip_field_mesh(nr=8, nz=8)
    m["er"] = np.zeros_like(m["er"])
    m["ez"] = np.full_like(m["ez"], ez)
    return m


def _one_ion(z0=1e-7, vz=500.0, t_inj=0.0):
    return {"x": np.array([2e-8]), "y": np.array([-1e-8]), "z": np.array([z0]),
            "vx": np.array([10.0]), "vy": np.array([-5.0]), "vz": np.array([vz]),
            "q": np.array([1.0]), "status": np.array([1], np.int32),
            "species": np.array([0], np.int32), "t_inj": np.array([t_inj])}


def test_one_level_is_leapfrog_dkd():
    m = synth.tip_field_mesh()
    d = synth.cone(600, seed=51)
    dt = 1e-12
    d["t_inj"] = np.sort(np.random.default_rng(3).uniform(0, 5 * dt, 600))
    kw = dict(mesh=m, bc_space=(0, 0, 9e-6, 2e-7), bc_time=6 * dt)
    a, fa = oracle.leapfrog_run_f64(d, 0.0, dt, 8, **kw)
    b, lev, fb = oracle.leapfrog_substeps_f64(d, 0.0, dt, 8, [NONE], **kw)
    for k in ("x", "y", "z", "vx", "vy", "vz", "status"):
        np.testing.assert_array_equal(a[k], b[k])
    np.testing.assert_array_equal(fa, fb)
    assert np.all(lev == 0)


@pytest.mark.parametrize("level", [0, 1, 2])
def test_uniform_field_parabola_at_every_level(level):
    m = _uniform_mesh()
    d = _one_ion()
    dt, k = 1e-12, 7
    s, lev, _ = oracle.leapfrog_substeps_f64(d, 0.0, dt, k, [NONE] * 3, level0=[level], mesh=m)
    a = oracle.lib.orc_e_over_amu() / 111.0 * m["E0"]
    T = dt * k
    assert lev[0] == level
    np.testing.assert_allclose(s["z"][0], d["z"][0] + d["vz"][0] * T + 0.5 * a * T * T, rtol=1e-13)
    np.testing.assert_allclose(s["vz"][0], d["vz"][0] + a * T, rtol=1e-13)
    np.testing.assert_allclose(s["x"][0], d["x"][0] + d["vx"][0] * T, rtol=1e-13)


def test_kick_window_of_an_ion_injected_mid_step():
    """Two levels (dt, dt/3), ion at level 0 injected at dt/2.  The six drifts
    of dt/6 end at dt/6 ... dt; it moves in the last three only.  Its one kick
    is the middle finest kick (top level 0), issued at tempVelTime = dt/3, so
    the reference clamps its window to min(dt/3 + dt - dt/2, dt) = 5 dt/6
    (plain DKD would give dt/2); the third kick (top level 1) leaves it alone."""
    m = _uniform_mesh()
    dt = 1e-12
    d = _one_ion(t_inj=0.5 * dt)
    s, _, _ = oracle.leapfrog_substeps_f64(d, 0.0, dt, 1, [NONE, NONE], level0=[0], mesh=m)
    a = oracle.lib.orc_e_over_amu() / 111.0 * m["E0"]
    v1 = d["vz"][0] + a * 5.0 * dt / 6.0
    np.testing.assert_allclose(s["vz"][0], v1, rtol=1e-14)
    np.testing.assert_allclose(s["z"][0], d["z"][0] + v1 * 0.5 * dt, rtol=1e-14)
    assert s["status"][0] == 2  # set by the last kick's setParticleStatus (level 0 < top 1)


SPACE = [(1, (0, 0, 0, 0), 0.0), (1, (0, 0, 2e-6, 0), 0.0), (1, (0, 0, 5e-7, 0), 0.0)]
TIME = [(2, (0, 0, 0, 0), 0.0), (2, (0, 0, 0, 0), 1e-11), (2, (0, 0, 0, 0), 2e-12)]


@pytest.mark.parametrize("z,expect", [(1e-7, 2), (1e-6, 1), (3e-6, 0), (-1e-6, 1)])
def test_spatial_levels(z, expect):
    assert oracle.substep_level_of(SPACE, 0, 2, 1e-8, 1e-8, z, 0.0, 0.0) == expect


def test_spatial_levels_from_an_inner_level():
    # crossoverParticles(1): only levels 1..2 are scanned; the x, y and rho axes
    # have no limit on them, so the ion stays at 1 however far it is
    assert oracle.substep_level_of(SPACE, 1, 2, 0.0, 0.0, 3e-6, 0.0, 0.0) == 1
    assert oracle.substep_level_of(SPACE, 1, 1, 0.0, 0.0, 1e-7, 0.0, 0.0) == 2


@pytest.mark.parametrize("age,expect", [(1e-12, 2), (5e-12, 1), (2e-11, 0)])
def test_temporal_levels(age, expect):
    assert oracle.substep_level_of(TIME, 0, 2, 0, 0, 0, 1e-9 + age, 1e-9) == expect


def test_space_and_time_take_the_finer_level():
    both = [(3, (0, 0, 0, 0), 0.0), (3, (0, 0, 2e-6, 0), 1e-11), (3, (0, 0, 5e-7, 0), 2e-12)]
    # near in space (level 2), old in time (level 0): the finer, 2
    assert oracle.substep_level_of(both, 0, 0, 0, 0, 1e-7, 1.0, 0.0) == 2
    # far in space (0), young (2)
    assert oracle.substep_level_of(both, 0, 0, 0, 0, 5e-6, 1e-12, 0.0) == 2
    # none of the levels admits the ion on any axis: -1 (never kicked again)
    tight = [(1, (1e-6, 1e-6, 1e-6, 1e-6), 0.0), (1, (5e-7, 5e-7, 5e-7, 5e-7), 0.0)]
    assert oracle.substep_level_of(tight, 0, 1, 2e-6, 2e-6, 2e-6, 0.0, 0.0) == -1
    # type none keeps the level
    assert oracle.substep_level_of([NONE, NONE], 0, 1, 0, 0, 1.0, 0.0, 0.0) == 1
