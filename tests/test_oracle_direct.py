"""Pins the oracle's direct sums (CPU only).

The reference ships no tests or golden vectors for this path, so the oracle is
pinned by closed forms, invariants, a long-double brute force and the
reference's own index algorithms."""
import itertools
import struct
from collections import Counter

import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _system(pos, q, species, status=None, v=None):
    pos = np.asarray(pos, dtype=np.float64)
    n = pos.shape[0]
    return {"x": pos[:, 0].copy(), "y": pos[:, 1].copy(), "z": pos[:, 2].copy(),
            "q": np.asarray(q, np.float64), "species": np.asarray(species, np.int32),
            "status": np.ones(n, np.int32) if status is None else np.asarray(status, np.int32)}


def test_k_const_bit_pattern():
    # K_CONST on x86-64 g++ (float e*e over a long double denominator)
    assert float.hex(oracle.k_const()) == "0x1.1c8a2aa5ab9edp-3"
    assert oracle.softening() == 6.4e-19
    assert [oracle.species_mass(s) for s in (0, 1, 2, 4, 5, 6, 7)] == [111, 309, 507, 198, 87, 285, 483]


def test_two_body_closed_form():
    d = 3.7e-9
    sysd = _system([[0, 0, 0], [d, 0, 0]], [1.0, -1.0], [0, 5])
    a = oracle.direct_f64(sysd)
    K, eps = oracle.k_const(), 6.4e-19
    mag = K * d / (d * d + eps) ** 1.5
    # a_0 = K q0/m0 * q1 (r0 - r1)/|.|^3 : attraction toward +x
    np.testing.assert_allclose(a[0], [mag * (1.0 / 111) * (-1.0) * (-1.0), 0, 0], rtol=1e-15)
    np.testing.assert_allclose(a[1], [mag * (-1.0 / 87) * 1.0 * 1.0, 0, 0], rtol=1e-15)


def test_three_body_collinear_closed_form():
    xs = np.array([0.0, 2e-9, 7e-9])
    q = np.array([1.0, 1.0, -1.0])
    sp = np.array([0, 1, 5])
    sysd = _system(np.stack([xs, 0 * xs, 0 * xs], 1), q, sp)
    a = oracle.direct_f64(sysd)
    K, eps = oracle.k_const(), 6.4e-19
    for i in range(3):
        s = 0.0
        for j in range(3):
            if j != i:
                dx = xs[i] - xs[j]
                s += q[j] * dx / (dx * dx + eps) ** 1.5
        expect = K * q[i] / oracle.species_mass(sp[i]) * s
        assert abs(a[i, 0] - expect) <= 1e-14 * abs(expect)
        assert a[i, 1] == 0 and a[i, 2] == 0


def test_momentum_identity_and_long_double():
    d = synth.cone(400, seed=7)
    a = oracle.direct_f64(d)
    m = np.array([oracle.species_mass(s) for s in d["species"]])
    p = (a * m[:, None]).sum(axis=0)
    scale = np.abs(a * m[:, None]).sum(axis=0)
    assert np.all(np.abs(p) <= 1e-13 * scale)
    ald = oracle.direct_f64_ld(d)
    err = np.linalg.norm(a - ald, axis=1) / np.linalg.norm(ald, axis=1)
    assert err.max() < 1e-13


def test_ref_contract_matches_f64_on_float_positions_and_rounds_like_reference():
    d = synth.hemisphere(300, seed=3)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    # Same physics on the float positions widened to double: differs only by the
    # float rounding of far-pair displacements (relative ~1e-7 at most).
    dd = dict(d)
    for c, k in enumerate("xyz"):
        dd[k] = pos4[:, c].astype(np.float64)
    f64 = oracle.direct_f64(dd)
    err = np.linalg.norm(ref - f64, axis=1) / np.linalg.norm(f64, axis=1)
    assert 0 < err.max() < 1e-6
    # Single image: the reference's float accumulation equals float(ai).
    acc = oracle.direct_ref_accumulate(pos4, vel4, acc4)
    np.testing.assert_array_equal(acc[:, :3], ref.astype(np.float32))


def test_status_rules():
    d = synth.hemisphere(64, seed=5)
    d["status"][:8] = 0      # out of bounds: neither target nor source
    d["status"][8:16] = 2    # source only
    d["status"][16:20] = -1  # crashed
    a = oracle.direct_f64(d)
    assert np.all(a[:8] == 0) and np.all(a[8:20] == 0)
    keep = np.ones(64, bool)
    keep[:8] = False
    keep[16:20] = False
    sub = {k: v[keep] for k, v in d.items()}
    b = oracle.direct_f64(sub)
    np.testing.assert_allclose(a[keep][sub["status"] == 1], b[sub["status"] == 1], rtol=1e-14)


def test_quadrant_rotation_table():
    # quadrantRotationFunction gives cos/sin of k*pi/2 (auxiliary_functions.h:55-57)
    assert [oracle.lib.orc_quadrant_cos(k) for k in range(4)] == [1, 0, -1, 0]
    assert [oracle.lib.orc_quadrant_sin(k) for k in range(4)] == [0, 1, 0, -1]


def _rotated_copies(d, nsec):
    n = d["x"].shape[0]
    xs, ys, zs, st = [d["x"]], [d["y"]], [d["z"]], [d["status"]]
    for k in range(1, nsec):
        a = 2 * np.pi * k / nsec
        if (4 * k) % nsec == 0:
            c, s = [(1, 0), (0, 1), (-1, 0), (0, -1)][(4 * k // nsec) % 4]
        else:
            c, s = np.cos(a), np.sin(a)
        xs.append(c * d["x"] - s * d["y"])
        ys.append(s * d["x"] + c * d["y"])
        zs.append(d["z"])
        st.append(np.where(d["status"] == 1, 2, d["status"]))
    return {"x": np.concatenate(xs), "y": np.concatenate(ys), "z": np.concatenate(zs),
            "q": np.tile(d["q"], nsec), "species": np.tile(d["species"], nsec),
            "status": np.concatenate(st)}, n


@pytest.mark.parametrize("nsec", [4, 6])
def test_rotational_equals_explicit_copies(nsec):
    d = synth.cone(200, seed=11)
    rot = oracle.direct_f64(d, sym="rotational", nsec=nsec)
    ext, n = _rotated_copies(d, nsec)
    plain = oracle.direct_f64(ext, i0=0, i1=n)
    err = np.linalg.norm(rot - plain, axis=1) / np.linalg.norm(plain, axis=1)
    assert err.max() < 1e-12


def test_rotational_ref_4fold_equals_explicit_float_copies():
    d = synth.hemisphere(150, seed=12)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    rot = oracle.direct_ref(pos4, vel4, acc4, sym="rotational", nsec=4)
    # 90-degree rotations are exact in float: build the copies explicitly.
    n = pos4.shape[0]
    P = [pos4]
    for (c, s) in [(0, 1), (-1, 0), (0, -1)]:
        p = pos4.copy()
        p[:, 0] = c * pos4[:, 0] - s * pos4[:, 1]
        p[:, 1] = s * pos4[:, 0] + c * pos4[:, 1]
        P.append(p)
    P = np.vstack(P)
    V = np.vstack([vel4] * 4)
    A = np.vstack([acc4] + [np.where(acc4 == 1, 2, acc4)] * 3)
    A[:n, 3] = acc4[:, 3]
    plain = oracle.direct_ref(P, V, A, i0=0, i1=n)
    err = np.linalg.norm(rot - plain, axis=1) / np.linalg.norm(plain, axis=1)
    assert err.max() < 1e-13


def test_translational_equals_explicit_copies():
    d = synth.cell_fill(120, seed=13)
    lat = synth.full_lattice(1, (1, 1, 1))
    dom = (2e-6, 2e-6, 2e-6)
    tr = oracle.direct_f64(d, sym="translational", lattice=lat, domain=dom)
    n = d["x"].shape[0]
    ext = {k: [] for k in ("x", "y", "z", "q", "species", "status")}
    # put the zero image first so rows 0..n are the real cell
    order = sorted(range(len(lat)), key=lambda k: tuple(abs(v) for v in lat[k]) != (0, 0, 0))
    for k in order:
        L = lat[k]
        ext["x"].append(d["x"] + L[0] * dom[0])
        ext["y"].append(d["y"] + L[1] * dom[1])
        ext["z"].append(d["z"] + L[2] * dom[2])
        ext["q"].append(d["q"])
        ext["species"].append(d["species"])
        ext["status"].append(d["status"] if not L.any() else np.full(n, 2, np.int32))
    ext = {k: np.concatenate(v) for k, v in ext.items()}
    plain = oracle.direct_f64(ext, i0=0, i1=n)
    err = np.linalg.norm(tr - plain, axis=1) / np.linalg.norm(plain, axis=1)
    assert err.max() < 1e-11


@pytest.mark.parametrize("box,axes", [(3, (1, 1, 1)), (5, (1, 1, 1)), (9, (1, 1, 0)),
                                      (3, (0, 1, 1)), (5, (1, 0, 0)), (5, (0, 0, 1))])
def test_reference_image_enumeration_is_the_full_block(box, axes):
    imgs = oracle.reference_lattice(box, axes)
    full = synth.full_lattice(box // 2, axes)
    assert Counter(map(tuple, imgs)) == Counter(map(tuple, full))


def test_rel_l2_hand_cases():
    ref = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 0]])
    assert oracle.rel_l2(ref, ref)[0] == 0.0
    l2, mx, k = oracle.rel_l2(ref * 1.001, ref)
    assert abs(l2 - 0.001 * np.sqrt(2 / 3)) < 1e-12 and k in (0, 1)
    assert oracle.rel_l2(ref[:0], ref[:0])[0] == 0.0
