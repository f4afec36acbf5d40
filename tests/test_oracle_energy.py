"""Pins the oracle's energy bookkeeping (CPU only) with closed forms: kinetic
and pair potential energy of a few ions, and the field energy of uniform
axial and radial fields along the reference's z-then-radial path."""
import numpy as np

import oracle
from solveit_hip import synth

AMU = float(np.float32(1.66053904e-27))
E = float(np.float32(1.6021766e-19))


def _ions(x, q, v, species, status=None):
    x = np.asarray(x, float)
    v = np.asarray(v, float)
    n = len(q)
    return {"x": x[:, 0], "y": x[:, 1], "z": x[:, 2], "q": np.asarray(q, float),
            "vx": v[:, 0], "vy": v[:, 1], "vz": v[:, 2],
            "species": np.asarray(species, np.int32),
            "status": np.ones(n, np.int32) if status is None else np.asarray(status, np.int32)}


def test_kinetic_and_potential_closed_form():
    x = [[0, 0, 0], [3e-9, 0, 0], [0, 4e-9, 0]]
    v = [[100.0, 0, 0], [0, 200.0, 0], [0, 0, -300.0]]
    d = _ions(x, [1, 1, -1], v, [0, 1, 5])
    ke, pe = oracle.energy_f64(d)
    m = [111.0, 309.0, 87.0]
    ke_ref = sum(0.5 * AMU * mi * np.dot(vi, vi) for mi, vi in zip(m, np.asarray(v)))
    ke2 = oracle.k_const() * AMU
    pe_ref = 0.0
    for i in range(3):
        for j in range(i + 1, 3):
            r2 = np.sum((np.asarray(x[i]) - np.asarray(x[j])) ** 2)
            pe_ref += ke2 * d["q"][i] * d["q"][j] / np.sqrt(r2 + 6.4e-19)
    assert abs(ke - ke_ref) <= 1e-14 * ke_ref
    assert abs(pe - pe_ref) <= 1e-14 * abs(pe_ref)
    # an ion with status 0 drops out of both sums
    d0 = dict(d)
    d0["status"] = np.array([1, 1, 0], np.int32)
    ke0, pe0 = oracle.energy_f64(d0)
    assert abs(ke0 - sum(0.5 * AMU * m[i] * np.dot(v[i], v[i]) for i in range(2))) < 1e-14 * ke0
    r2 = 9e-18
    assert abs(pe0 - ke2 / np.sqrt(r2 + 6.4e-19)) < 1e-14 * abs(pe0)


def _uniform_mesh(er, ez):
    m = dict(synth.tip_field_mesh())
    m["er"] = np.full_like(m["er"], er)
    m["ez"] = np.full_like(m["ez"], ez)
    return m


def test_field_energy_uniform_axial():
    m = _uniform_mesh(0.0, 1.0)
    rng = np.random.default_rng(3)
    n = 50
    r0 = m["r0"]
    p0 = (rng.uniform(0, 5, n) * r0 * 0.1, rng.uniform(0, 5, n) * r0 * 0.1, rng.uniform(1, 60, n) * r0)
    p1 = (p0[0] + rng.normal(0, 1e-8, n), p0[1] + rng.normal(0, 1e-8, n),
          p0[2] + rng.uniform(-1e-7, 1e-6, n))
    q = np.where(rng.random(n) < 0.1, -1.0, 1.0)
    fe = oracle.field_energy_f64(m, p0, p1, q, nsteps=7)
    expect = -E * q * m["E0"] * (p1[2] - p0[2])
    np.testing.assert_allclose(fe, expect, rtol=1e-12)


def test_field_energy_uniform_radial():
    """E = E0 (x, y, 0) / rho: along a ray away from the axis the radial leg
    gives -e q E0 (rho1 - rho0); the axial leg gives nothing."""
    m = _uniform_mesh(1.0, 0.0)
    n = 20
    rng = np.random.default_rng(4)
    phi = rng.uniform(0, 2 * np.pi, n)
    rho0 = rng.uniform(0.1, 3.0, n) * 1e-6
    rho1 = rho0 * rng.uniform(1.0, 2.0, n)
    z0 = rng.uniform(1, 50, n) * m["r0"]
    z1 = z0 + rng.uniform(0, 1e-6, n)
    p0 = (rho0 * np.cos(phi), rho0 * np.sin(phi), z0)
    p1 = (rho1 * np.cos(phi), rho1 * np.sin(phi), z1)
    q = np.ones(n)
    fe = oracle.field_energy_f64(m, p0, p1, q, nsteps=np.full(n, 5, np.int32))
    np.testing.assert_allclose(fe, -E * m["E0"] * (rho1 - rho0), rtol=1e-10)
