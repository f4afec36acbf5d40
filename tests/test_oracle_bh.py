"""Pins the oracle's Barnes-Hut restatement (CPU only): closed forms for two
ions and for the reference's centre-of-charge rule, theta -> 0 against the
direct sum, accuracy falling monotonically with theta, and the symmetry
variants at theta = 0 against the direct sums with images."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth


def _aos(x, q, status=None, species=None):
    n = len(q)
    pos4 = np.zeros((n, 4), np.float32)
    pos4[:, :3] = np.asarray(x, np.float32)
    pos4[:, 3] = q
    vel4 = np.zeros((n, 4), np.float32)
    vel4[:, 3] = 0 if species is None else species
    acc4 = np.zeros((n, 4), np.float32)
    acc4[:, 3] = 1 if status is None else status
    return pos4, vel4, acc4


def test_two_ions_closed_form():
    x = [[0.0, 0.0, 0.0], [3e-9, 4e-9, 0.0]]
    pos4, vel4, acc4 = _aos(x, [1.0, 1.0])
    out, ncell = oracle.bh_ref(pos4, vel4, acc4, theta=0.5)
    kc = oracle.k_const() / 111.0
    r = np.array(x[0], np.float32) - np.array(x[1], np.float32)
    d = np.sqrt(float(np.float32(r @ r)) + 6.4e-19)
    expect = kc * r / d ** 3
    np.testing.assert_allclose(out[0], expect, rtol=1e-6)
    np.testing.assert_allclose(out[1], -expect, rtol=1e-6)
    assert ncell == 9  # root split once into 8 children


def test_centre_of_charge_rule():
    """Root cube [0, 8 nm]^3: its child 0 holds a +1/-1 pair near the origin
    (net charge 0) and a +1 ion at b = 3 nm; the target sits at 8 nm.  The
    reference weights child centres by |net charge| over sum |q|, so child 0's
    centre is (1 x b) / 3; at theta = 0.7 the target opens the root and
    accepts child 0 as one monopole (its own leaf is skipped)."""
    b = np.array([3e-9, 3e-9, 3e-9])
    far = np.array([8e-9, 8e-9, 8e-9])
    x = [np.zeros(3), np.full(3, 1e-10), b, far]
    pos4, vel4, acc4 = _aos(x, [1.0, -1.0, 1.0, 1.0], status=[2, 2, 2, 1])
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.7)
    c = (np.float32(1.0) * b.astype(np.float32)) / np.float32(3.0)
    r = far.astype(np.float32) - c
    d = np.float32(np.sqrt(float(np.float32(r @ r)) + 6.4e-19))
    expect = oracle.k_const() / 111.0 * r / np.float64(d) ** 3
    np.testing.assert_allclose(out[3], expect, rtol=1e-6)
    # and not the |q|-weighted mean of the three ions
    wrong = far - (0 + 1e-10 + 3e-9) / 3.0
    assert abs(out[3][0] - oracle.k_const() / 111.0 * wrong[0] / np.linalg.norm(wrong) ** 3) > 1e-3 * abs(out[3][0])


def test_theta_zero_is_direct():
    d = synth.cone(2500, seed=31)
    d["status"][::17] = 2
    d["status"][::41] = 0
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    m = d["status"] == 1
    assert np.all(out[~m] == 0)
    l2, mx, _ = oracle.rel_l2(out, ref, m)
    assert l2 < 5e-5 and mx < 1e-3  # float accumulation of the same pair terms


def test_accuracy_monotone_in_theta():
    d = synth.cone(3000, seed=32)
    d["q"][:] = 1.0  # one sign: the monopole centre is the |q|-weighted mean
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    ref = oracle.direct_f64(d)
    errs = [oracle.rel_l2(oracle.bh_ref(pos4, vel4, acc4, theta=t)[0], ref)[0]
            for t in (0.2, 0.4, 0.6)]
    assert errs[0] < errs[1] < errs[2] < 2e-2


@pytest.mark.parametrize("sym", ["rotational", "translational"])
def test_symmetry_theta_zero_is_direct(sym):
    if sym == "rotational":
        d = synth.cone(1200, seed=33)
        kw = dict(sym="rotational")
    else:
        d = synth.cell_fill(800, seed=34)
        kw = dict(sym="translational", lattice=synth.full_lattice(1, (1, 1, 1)),
                  domain=(2e-6, 2e-6, 2e-6))
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0, **kw)
    ref = oracle.direct_ref(pos4, vel4, acc4, **kw)
    l2, mx, _ = oracle.rel_l2(out, ref, d["status"] == 1)
    assert l2 < 5e-5


S = 2.0 ** -20  # 9.5e-7 m: grid arithmetic in units of S stays exact in float


def _box_cloud(n, ext, seed, corner_status=1):
    """n ions in [0, ext] (units of S) with two ions on opposite corners, so
    the bounding box is exactly [0, ext]."""
    rng = np.random.default_rng(seed)
    ext = np.asarray(ext, np.float64)
    x = rng.uniform(0.05, 0.95, size=(n, 3)) * ext * S
    x[0] = 0.0
    x[1] = ext * S
    q = np.where(rng.random(n) < 0.1, -1.0, 1.0)
    st = np.ones(n)
    st[:2] = corner_status
    return _aos(x, q, status=st)


@pytest.mark.parametrize("n,nd,size", [
    # 1 x 1 x 4 box: cells of side 1 give 4 cells; >= 100 ions per cell needs n >= 397
    (400, (1, 1, 4), 1.0),
    # too few ions: side doubled to 2 (ratio 0.5 rounds up to 1 cell, 4 / 2 = 2 cells)
    (300, (1, 1, 2), 2.0),
    # doubled again to 4: one cell of the longest extent
    (150, (1, 1, 1), 4.0),
])
def test_layout_hand_cases(n, nd, size):
    """computeDomainBounds + computeWorldLimits on a 1 x 1 x 4 box (units of S),
    derived by hand: the cell counts per axis, the cell side (size x (1 +
    0.01 / n_long), n_long the count along the long axis) and the grid centred
    on the box."""
    pos4, vel4, acc4 = _box_cloud(n, (1, 1, 4), seed=n)
    roots, dims, got_nd, ns = oracle.bh_layout(pos4, acc4)
    assert ns == n
    assert tuple(got_nd) == nd
    side = np.float32(size * S) * (np.float32(1) + np.float32(0.01) / np.float32(nd[2]))
    np.testing.assert_array_equal(dims, np.full(3, side, np.float32))
    ext = np.array([1, 1, 4], np.float32) * np.float32(S)
    expect_roots = (ext + np.float32(0) - side * np.asarray(nd, np.float32)) / np.float32(2)
    np.testing.assert_array_equal(roots, expect_roots.astype(np.float32))


def test_layout_forced_counts_merge():
    """setBoxDimVals(2, 2, 4) on the 1 x 1 x 4 box: side max(1/2, 1/2, 4/4) = 1,
    scaled by 1 + 0.01 / 4; computeWorldLimits then merges 2 x 2 x 2 cells
    because the smallest count is 2, giving 1 x 1 x 2 cells of twice the side."""
    pos4, vel4, acc4 = _box_cloud(400, (1, 1, 4), seed=7)
    roots, dims, nd, _ = oracle.bh_layout(pos4, acc4, box_dims=(2, 2, 4))
    assert tuple(nd) == (1, 1, 2)
    side = np.float32(S) * (np.float32(1) + np.float32(0.01) / np.float32(4))
    np.testing.assert_array_equal(dims, np.full(3, np.float32(2) * side, np.float32))


def test_layout_prism_and_unmatched_root():
    """Rectangular-prism root: one cell equal to the bounding box.  The ion on
    the far corner gives (pos - roots) / dims = 1 on every axis, which matches
    no cell; the reference then carries it to the last cell of the chain."""
    pos4, vel4, acc4 = _box_cloud(50, (1, 1, 4), seed=8)
    roots, dims, nd, _ = oracle.bh_layout(pos4, acc4, prism=True)
    assert tuple(nd) == (1, 1, 1)
    np.testing.assert_array_equal(roots, np.zeros(3, np.float32))
    np.testing.assert_array_equal(dims, np.array([1, 1, 4], np.float32) * np.float32(S))
    out, _, info = oracle.bh_ref(pos4, vel4, acc4, theta=0.5, prism=True, info=True)
    assert info["root_cells"] == 1 and info["unmatched_root"] == 1
    ref = oracle.direct_ref(pos4, vel4, acc4)
    out0, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0, prism=True)
    assert oracle.rel_l2(out0, ref)[0] < 5e-5


def test_layout_seeded_by_first_ion_of_range():
    """The bounds start from the first ion of the range whatever its status
    (domain_bounds_functions.cpp:139-150): an excluded far ion at index 0
    widens the grid as a source there would, but adds no force."""
    pos4, vel4, acc4 = _box_cloud(400, (1, 1, 4), seed=9)
    far = np.array([0.5, 0.5, 12.0]) * S
    pos4 = np.vstack([np.r_[far, 1.0][None].astype(np.float32), pos4])
    vel4 = np.vstack([vel4[:1], vel4])
    acc4 = np.vstack([acc4[:1], acc4])
    acc4[0, 3] = 0  # out: neither target nor source
    lay_out = oracle.bh_layout(pos4, acc4)
    acc_src = acc4.copy()
    acc_src[0, 3] = 2
    lay_src = oracle.bh_layout(pos4, acc_src)
    lay_none = oracle.bh_layout(pos4, acc4, i0=1)
    for a, b in zip(lay_out[:3], lay_src[:3]):
        np.testing.assert_array_equal(a, b)
    top = lambda lay: lay[0][2] + lay[2][2] * lay[1][2]
    assert top(lay_out) >= np.float32(far[2]) > top(lay_none)
    out, _ = oracle.bh_ref(pos4, vel4, acc4, theta=0.0)
    ref = oracle.direct_ref(pos4, vel4, acc4)
    assert oracle.rel_l2(out, ref, acc4[:, 3] == 1)[0] < 5e-5


def test_cone_root_grid_covers_sources():
    """The cone plume is about 10 um long and 2-3 um wide: the reference lays
    several cubic root cells along it, and every source lies inside the grid."""
    d = synth.cone(6000, seed=42)
    pos4, vel4, acc4 = synth.to_reference_aos(d)
    roots, dims, nd, ns = oracle.bh_layout(pos4, acc4)
    assert nd.prod() > 1 and nd[2] == nd.max()
    m = (acc4[:, 3] == 1) | (acc4[:, 3] == 2)
    p = pos4[m, :3]
    assert np.all(p >= roots) and np.all(p <= roots + nd * dims)
    _, _, info = oracle.bh_ref(pos4, vel4, acc4, theta=0.5, info=True)
    assert info["root_cells"] == nd.prod() and info["buckets"] == 0
    assert info["max_depth"] < 42
