"""Pins the oracle's background field (CPU only).

The reference's own pwl_interp_2d_scattered.c and InterpolationScheme class are
compiled from /root/reference by oracle/ref.mk (where the tree is present) and
compared bit for bit; the interpolation is also checked against fields it must
reproduce exactly (linear in (r, z))."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from solveit_hip import synth


@pytest.fixture(scope="module")
def mesh():
    return synth.tip_field_mesh()


def _inside_points(rng, n, hi=140.0):
    # strictly inside the hull, biased toward the fine cells near the tip
    r = hi * rng.random(n) ** 3 * 0.999 + 1e-9
    z = hi * rng.random(n) ** 3 * 0.999 + 1e-9
    return np.stack([r, z], 1)


def _prime_start(mesh, p0):
    e, t, _, _ = oracle.tri_search(mesh, p0, start=-1)
    assert e == 0
    return t


def test_walk_matches_reference_pwl_bitwise(mesh):
    L = oracle.ref_pwl_lib()
    if L is None:
        pytest.skip("reference pwl build (oracle/_ref) not present")
    libc = ctypes.CDLL(None)
    libc.free.argtypes = [ctypes.c_void_p]
    m = {k: np.ascontiguousarray(mesh[k]) for k in ("xy", "er", "ez", "tri", "nbr")}
    nd, nt = m["er"].shape[0], m["tri"].shape[0]

    def ref_value(z, p):
        pp = np.ascontiguousarray(p, dtype=np.float64)
        ptr = L.pwl_interp_2d_scattered_value(nd, m["xy"].ctypes.data, z.ctypes.data, nt,
                                              m["tri"].ctypes.data, m["nbr"].ctypes.data, 1,
                                              pp.ctypes.data)
        v = ctypes.cast(ptr, ctypes.POINTER(ctypes.c_double))[0]
        libc.free(ptr)
        return v

    rng = np.random.default_rng(1)
    pts = _inside_points(rng, 400)
    # prime the reference's static warm start with a point strictly inside a triangle
    p0 = np.array([0.37, 0.41])
    ref_value(m["er"], p0)
    t = _prime_start(mesh, p0)
    for p in pts:
        vr_ref = ref_value(m["er"], p)
        vz_ref = ref_value(m["ez"], p)
        e, t, bc, _ = oracle.tri_search(mesh, p, start=t)
        a, b, c = m["tri"][t]
        vr = bc[0] * m["er"][a] + bc[1] * m["er"][b] + bc[2] * m["er"][c]
        vz = bc[0] * m["ez"][a] + bc[1] * m["ez"][b] + bc[2] * m["ez"][c]
        assert e == 0
        assert vr == vr_ref and vz == vz_ref


def test_efield_matches_reference_interpolation_scheme(mesh, tmp_path):
    L = oracle.ref_efield_lib()
    if L is None:
        pytest.skip("reference InterpolationScheme build (oracle/_ref) not present")
    fields, tris = synth.write_mesh_csv(mesh, str(tmp_path))
    r0, E0 = np.float32(mesh["r0"]), np.float32(mesh["E0"])
    L.ref_efield_init((str(tmp_path) + "/").encode(), fields.encode(), tris.encode(), r0, E0)
    rng = np.random.default_rng(2)
    d = synth.cone(300, seed=3)
    pos = np.stack([d["x"], d["y"], d["z"]], 1).astype(np.float32)
    pos[:5, 2] *= -1          # below the tip: gated to zero
    pos[5:8, 0] = 0.0         # on the axis planes
    pos[8, :2] = 0.0
    out = np.zeros(3, np.float32)
    # prime both warm starts with an interior point
    L.ref_efield(np.float32(1e-6), np.float32(0.5e-6), np.float32(2e-6), out.ctypes.data)
    t = _prime_start(mesh, np.array([np.hypot(np.float32(1e-6) / r0, np.float32(0.5e-6) / r0),
                                     np.float32(2e-6) / r0]))
    mine = np.zeros(3, np.float32)
    ti = ctypes.c_int(t)
    mm = {k: np.ascontiguousarray(mesh[k]) for k in ("xy", "er", "ez")}
    mi = {k: np.ascontiguousarray(mesh[k], dtype=np.int32) for k in ("tri", "nbr")}
    for p in pos:
        L.ref_efield(p[0], p[1], p[2], out.ctypes.data)
        oracle.lib.orc_efield_ref(p[0], p[1], p[2], r0, E0, mm["xy"].ctypes.data,
                                  mi["tri"].shape[0], mi["tri"].ctypes.data, mi["nbr"].ctypes.data,
                                  mm["er"].ctypes.data, mm["ez"].ctypes.data,
                                  ctypes.addressof(ti), mine.ctypes.data)
        np.testing.assert_array_equal(mine, out)
    assert rng is not None


def test_linear_field_is_reproduced(mesh):
    lin = dict(mesh)
    lin["er"] = 0.3 + 0.01 * mesh["xy"][:, 0] - 0.002 * mesh["xy"][:, 1]
    lin["ez"] = 1.0 - 0.004 * mesh["xy"][:, 1] + 0.001 * mesh["xy"][:, 0]
    rng = np.random.default_rng(3)
    pts = _inside_points(rng, 500) * mesh["r0"]
    phi = rng.uniform(0, 2 * np.pi, pts.shape[0])
    x, y, z = pts[:, 0] * np.cos(phi), pts[:, 0] * np.sin(phi), pts[:, 1]
    E, edge, _ = oracle.efield_f64(lin, x, y, z)
    rn, zn = pts[:, 0] / mesh["r0"], pts[:, 1] / mesh["r0"]
    er = 0.3 + 0.01 * rn - 0.002 * zn
    ez = 1.0 - 0.004 * zn + 0.001 * rn
    expect = mesh["E0"] * np.stack([er * np.cos(phi), er * np.sin(phi), ez], 1)
    assert np.all(edge == 0)
    np.testing.assert_allclose(E, expect, rtol=1e-11, atol=1e-9 * mesh["E0"])


def test_gate_and_axis(mesh):
    r0 = mesh["r0"]
    E, _, _ = oracle.efield_f64(mesh, [0.0, 1e-7, 0.0], [0.0, 0.0, 0.0],
                                [-1e-9, 141 * r0, 0.0])
    assert np.all(E[:2] == 0)
    # on the axis the field is (0, 0, E0 * ez(0, 0)) = (0, 0, E0)
    np.testing.assert_allclose(E[2], [0, 0, mesh["E0"]], rtol=1e-12)


def test_field_mesh_neighbours_are_consistent(mesh):
    tri, nbr = mesh["tri"], mesh["nbr"]
    for t in range(0, tri.shape[0], 389):
        for k in range(3):
            u = nbr[t, k]
            if u < 0:
                continue
            shared = set(tri[t]) - {tri[t, k]}
            kk = list(nbr[u]).index(t)
            assert set(tri[u]) - {tri[u, kk]} == shared


def test_analytic_field_matches_finite_difference_potential():
    # E = -grad(phi) of the hyperboloid potential, up to the apex normalisation
    a = 200.0
    k = 1.0 / a
    eta0 = (-k + np.sqrt(k * k + 4)) / 2
    d = a * eta0

    def phi(rho, z):
        Z = d - z
        r1 = np.hypot(rho, Z - a)
        r2 = np.hypot(rho, Z + a)
        eta = (r2 - r1) / (2 * a)
        return np.log((1 + eta) / (1 - eta))

    rho, z, h = np.array([0.3, 2.0, 50.0]), np.array([0.5, 10.0, 100.0]), 1e-5
    er, ez = synth.hyperboloid_field(rho, z)
    _, ez_apex = synth.hyperboloid_field(np.array([0.0]), np.array([0.0]))
    norm = -(phi(0.0, h) - phi(0.0, -h)) / (2 * h)
    fr = -(phi(rho + h, z) - phi(rho - h, z)) / (2 * h) / norm
    fz = -(phi(rho, z + h) - phi(rho, z - h)) / (2 * h) / norm
    np.testing.assert_allclose(er, fr, rtol=1e-6, atol=1e-12)
    np.testing.assert_allclose(ez, fz, rtol=1e-6)
    assert abs(ez_apex[0] - 1.0) < 1e-15
