"""Pins of the oracle's restatement of the reference FMM's own operators
(oracle/fmm_sph.c: MultipoleMethodConstants::precalc, rotation, p2m / m2m /
m2l / l2l / l2p kernels of fmm_cpukernel.cpp) by what the mathematics fixes:
closed-form tables, rotations that undo each other, translations that are
exact (M2M, L2L), convergence to the exact Coulomb field, and the tree rules
of setDomainSize / setOptimumLevel on hand cases."""
import math

import numpy as np
import pytest

import oracle


def _cube(n, side, seed, zscale=1.0):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, side, (n, 3))
    pos[:, 2] *= zscale
    pos[:, 3] = np.where(rng.random(n) < 0.05, -1.0, 1.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = rng.integers(0, 2, n)  # monomer / dimer
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    return pos, vel, acc


def test_tables_closed_forms():
    """factorial = sqrt((n-|m|)!/(n+|m|)!), anm = (-1)^n / sqrt((n-m)!(n+m)!)
    (both stored in float), and Ynm on the z axis: 1 for m = 0, 0 otherwise."""
    for n in range(20):
        for m in range(-n, n + 1):
            nm = n * n + n + m
            f = math.sqrt(math.factorial(n - abs(m)) / math.factorial(n + abs(m)))
            a = (-1) ** n / math.sqrt(math.factorial(n - m) * math.factorial(n + m))
            assert oracle.sph_table("factorial", nm) == float(np.float32(f))
            assert oracle.sph_table("anm", nm) == float(np.float32(a))
            assert oracle.sph_table("Ynm", nm) == pytest.approx(1.0 if m == 0 else 0.0, abs=1e-7)


def test_Anm_from_anm():
    """Anm[j,k; n,k] = (-1)^(j+k) anm[n,k] anm[j,k] / anm[j+n,0] (float)."""
    for j in range(10):
        for k in range(-j, j + 1):
            for n in range(abs(k), 10):
                jk, nk, jn = j * j + j + k, n * n + n + k, (j + n) * (j + n) + j + n
                want = (-1) ** (j + k) * oracle.sph_table("anm", nk) * oracle.sph_table("anm", jk) / \
                    oracle.sph_table("anm", jn)
                assert oracle.sph_table("Anm", jk * 100 + nk) == float(np.float32(want))


def _multipole(p, seed, h=1e-8):
    rng = np.random.default_rng(seed)
    xyzq = np.c_[rng.uniform(-0.5, 0.5, (20, 3)) * h, rng.choice([-1.0, 1.0], 20)]
    return oracle.sph_p2m(p, xyzq, h)


@pytest.mark.parametrize("d", [(1, 0, 0), (1, 1, 1), (2, -1, 3), (-3, 1, 0), (1, -1, -1), (0, 1, 2),
                               (-5, 4, 3)])
def test_rotation_round_trip(d):
    """The backward rotation undoes the forward one on a multipole of real charges."""
    p = 10
    C = _multipole(p, 1)
    R = oracle.sph_rotate(p, d, C)
    B = oracle.sph_rotate(p, d, R, backward=True)
    assert np.max(np.abs(B - C)) <= 1e-14 * np.max(np.abs(C))
    assert np.max(np.abs(R - C)) > 1e-3 * np.max(np.abs(C))  # the rotation does something


def test_rotation_preserves_degree_norms():
    """A rotation mixes orders within a degree only: sum over m of |C_n^m|^2
    with weights 1 (m = 0) and 2 (m > 0) is invariant, up to the normalisation
    factor of the reference's expansions, which is itself m-independent."""
    p = 10
    C = _multipole(p, 2)
    R = oracle.sph_rotate(p, (2, -1, 3), C)

    def norms(V):
        out = []
        for n in range(p):
            s = abs(V[n * (n + 1) // 2]) ** 2
            s += 2 * sum(abs(V[n * (n + 1) // 2 + m]) ** 2 for m in range(1, n + 1))
            out.append(s)
        return np.array(out)

    np.testing.assert_allclose(norms(R), norms(C), rtol=1e-12)


@pytest.mark.parametrize("octant", range(8))
def test_m2m_and_l2l_are_exact(octant):
    """M2M of a child's multipole equals the parent's multipole of the same
    charges, and L2L re-expands a local expansion without loss, up to the
    float storage of the tables and the float angles (below 5e-7)."""
    p, H = 10, 1e-6
    rng = np.random.default_rng(octant)
    cx, cy, cz = (octant >> 1) & 1, octant & 1, (octant >> 2) & 1
    cc = np.array([cx - 0.5, cy - 0.5, cz - 0.5]) * H / 2
    pts = rng.uniform(-0.5, 0.5, (10, 3)) * H / 2
    q = rng.choice([-1.0, 1.0], 10)
    Mp = oracle.sph_m2m(p, octant, oracle.sph_p2m(p, np.c_[pts, q], H / 2))
    Md = oracle.sph_p2m(p, np.c_[pts + cc, q], H)
    assert np.max(np.abs(Mp - Md)) <= 5e-7 * np.max(np.abs(Md))
    M = oracle.sph_p2m(p, [0.1 * H, 0, 0, 1.0], H)
    Lp = oracle.sph_m2l(p, (3, 0, 1), H, M)
    Lc = oracle.sph_l2l(p, octant, Lp)
    x = rng.uniform(-0.5, 0.5, 3) * H / 2
    f1 = oracle.sph_l2p(p, Lc, x, H / 2)
    f2 = oracle.sph_l2p(p, Lp, x + cc, H)
    assert np.linalg.norm(f1 - f2) <= 5e-7 * np.linalg.norm(f2)


def test_single_source_converges_to_coulomb():
    """P2M -> M2L -> L2P of one charge gives -q (r_t - r_s)/|r|^3 (the field the
    reference's L2P subtracts), converging geometrically in p."""
    h = 1e-6
    rng = np.random.default_rng(2)
    worst = {}
    for p in (4, 6, 8, 10):
        errs = []
        for t in range(12):
            src = rng.uniform(-0.5, 0.5, 3) * h
            d = np.array([3, 1, -2]) if t % 2 else np.array([2, 2, 2])
            off = rng.uniform(-0.5, 0.5, 3) * h
            L = oracle.sph_m2l(p, d, h, oracle.sph_p2m(p, [*src, 1.0], h))
            f = oracle.sph_l2p(p, L, off, h)
            r = (d * h + off) - src
            exact = r / np.linalg.norm(r) ** 3
            errs.append(np.linalg.norm(-f - exact) / np.linalg.norm(exact))
        worst[p] = max(errs)
    assert worst[4] > worst[6] > worst[8] > worst[10]
    assert worst[10] < 1e-5 and worst[4] < 5e-2


def test_fmm_converges_to_direct_sum():
    """On a 10 um cube (leaf boxes far larger than the softening length) the
    whole FMM converges to the direct sum as p grows."""
    pos, vel, acc = _cube(6000, 1e-5, 4)
    ref = oracle.direct_ref(pos, vel, acc)
    errs = []
    for p in (2, 4, 6, 8, 10):
        a, info = oracle.fmm_sph(pos, vel, acc, p=p, max_level=3)
        assert (info["nd0"], info["nd1"], info["nd2"], info["k"]) == (1, 1, 1, 0)
        errs.append(oracle.rel_l2(a, ref)[0])
    assert all(e1 > e2 for e1, e2 in zip(errs, errs[1:])), errs
    assert errs[-1] < 3e-5


def test_softening_floor_of_the_reference_expansion():
    """The reference expands plain 1/r and softens only the M2L centre distance,
    so on a dense 100 nm cube (12.5 nm leaves against a 0.8 nm softening
    length) the error stops falling with p (SURVEY F9's hemisphere floor)."""
    pos, vel, acc = _cube(8000, 1e-7, 5)
    ref = oracle.direct_ref(pos, vel, acc)
    e8 = oracle.rel_l2(oracle.fmm_sph(pos, vel, acc, p=8)[0], ref)[0]
    e10 = oracle.rel_l2(oracle.fmm_sph(pos, vel, acc, p=10)[0], ref)[0]
    assert e10 > 1e-4 and e10 > 0.7 * e8


def test_level_switch_and_domain_grid():
    """setOptimumLevel: 6000 sources in one domain -> maxLevel 2, 20000 -> 3;
    a 4:4:1 slab -> a 4 x 4 x 1 domain grid, two octree levels above it."""
    for n, want in ((6000, 2), (20000, 3)):
        pos, vel, acc = _cube(n, 1e-5, 6)
        _, info = oracle.fmm_sph(pos, vel, acc, p=2)
        assert info["max_level"] == want and info["sources"] == n
        assert info["leaves"] <= 8 ** want
    pos, vel, acc = _cube(6000, 1e-5, 7, zscale=0.25)
    bm, dims, nd, ns = oracle.domain_bounds(pos, acc)
    assert tuple(nd) == (4, 4, 1) and ns == 6000
    _, info = oracle.fmm_sph(pos, vel, acc, p=4, max_level=2)
    assert (info["nd0"], info["nd1"], info["nd2"], info["k"]) == (4, 4, 1, 2)


def test_flattened_multi_domain_converges():
    pos, vel, acc = _cube(6000, 1e-5, 7, zscale=0.25)
    ref = oracle.direct_ref(pos, vel, acc)
    e = [oracle.rel_l2(oracle.fmm_sph(pos, vel, acc, p=p, max_level=2)[0], ref)[0] for p in (4, 8, 10)]
    assert e[0] > e[1] > e[2] and e[2] < 5e-5


def test_status_rules_and_range():
    """Targets are status 1 in [i0, i1); status 2 ions are sources only; status 0
    ions and ions outside the range take no part."""
    pos, vel, acc = _cube(3000, 1e-5, 8)
    acc[::7, 3] = 2.0
    acc[::11, 3] = 0.0
    i0, i1 = 100, 2900
    a, _ = oracle.fmm_sph(pos, vel, acc, p=10, max_level=2, i0=i0, i1=i1)
    ref = oracle.direct_ref(pos, vel, acc, i0=i0, i1=i1, j0=i0, j1=i1)
    tgt = acc[i0:i1, 3] == 1.0
    assert np.all(a[~tgt] == 0.0)
    assert oracle.rel_l2(a[tgt], ref[tgt])[0] < 5e-5


# ---- rotational symmetry (updateAccelerationRotational) ----

def _quadrant_column(n, side, zlen, seed):
    """Ions in the first quadrant x, y in [0, side), z in [0, zlen)."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(0, side, n)
    pos[:, 1] = rng.uniform(0, side, n)
    pos[:, 2] = rng.uniform(0, zlen, n)
    pos[:, 3] = np.where(rng.random(n) < 0.05, -1.0, 1.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = rng.integers(0, 2, n)
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    return pos, vel, acc


def test_rotational_domain_bounds_hand_case():
    """computeDomainBoundsRotations: a square [-R, R]^2 about the axis, R the
    larger of the x and y maxima.  Sources with max x 3 um, max y 5 um and z in
    [1, 13] um: extents (10, 10, 12) um, the smallest (x) is the first cell size,
    y and z stay one cell (ratios 1 and 1.2), the size becomes the z extent,
    then x 1.01 (domainEps) and the grid is centred on the axis in x, y."""
    rng = np.random.default_rng(3)
    n = 400
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = rng.uniform(0.5e-6, 3e-6, n)
    pos[:, 1] = rng.uniform(0.5e-6, 5e-6, n)
    pos[:, 2] = rng.uniform(1e-6, 13e-6, n)
    pos[0, :3] = (1e-6, 1e-6, 1e-6)
    pos[1, :3] = (3e-6, 5e-6, 13e-6)
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    bm, dims, nd, ns = oracle.domain_bounds(pos, acc, rotational=True)
    f = np.float32
    R = f(5e-6)
    size = f(f(13e-6) - f(1e-6)) * f(f(1.0) + f(0.01) / f(1.0))
    assert tuple(nd) == (1, 1, 1) and ns == n
    assert dims[0] == size
    assert bm[0] == f((R + (-R) - size) / f(2)) == bm[1]
    assert bm[2] == f((f(13e-6) + f(1e-6) - size) / f(2))
    # the plain bounds of the same ions are not centred on the axis
    bm0, _, _, _ = oracle.domain_bounds(pos, acc)
    assert bm0[0] != bm[0]


@pytest.mark.parametrize("zlen,level,bar", [(1.5e-5, -2, 1e-5), (4e-5, -2, 1e-5), (1.2e-4, 3, 3e-5)])
def test_rotational_fmm_converges_to_rotational_direct(zlen, level, bar):
    """The rotational FMM (quadrant map in the P2P and M2L lists, i^(-mk) phase
    of imaged multipoles, four-image M2M into the domains) converges with p to
    the direct sum over the four quadrant images (direct_rotational).  The
    cases: one domain, two domains along z, and six (K = 3, where the M2M into
    the domains is a level of the tree)."""
    pos, vel, acc = _quadrant_column(6000 if zlen != 4e-5 else 12000, 1e-5, zlen, 1)
    ref = oracle.direct_ref(pos, vel, acc, sym="rotational", nsec=4)
    errs = []
    for p in (4, 8, 10):
        a, info = oracle.fmm_sph(pos, vel, acc, p=p, max_level=level, rotational=True)
        assert (info["nd0"], info["nd1"]) == (1, 1)
        errs.append(oracle.rel_l2(a, ref)[0])
    if zlen == 1.2e-4:
        assert info["nd2"] == 6 and info["k"] == 3
    assert errs[0] > errs[1] > errs[2] and errs[2] < bar, errs
    # without the images the plain FMM misses them entirely
    a0, _ = oracle.fmm_sph(pos, vel, acc, p=10, max_level=level)
    assert oracle.rel_l2(a0, ref)[0] > 0.1


def test_rotational_wide_grid_is_not_restated():
    """A quadrant cube has a 2 x 2 x 1 rotational grid, where the reference's
    quadrant map (domain-local x, y) does not hold."""
    pos, vel, acc = _quadrant_column(3000, 1e-5, 1e-5, 2)
    _, _, nd, _ = oracle.domain_bounds(pos, acc, rotational=True)
    assert tuple(nd) == (2, 2, 1)
    with pytest.raises(ValueError):
        oracle.fmm_sph(pos, vel, acc, p=4, rotational=True)


def test_rotational_status_rules_and_range():
    """Rotational: status-2 ions are sources only (their images too), status-0
    ions and ions outside [i0, i1) take no part, as in direct_rotational over
    the same range."""
    pos, vel, acc = _quadrant_column(5000, 1e-5, 1.5e-5, 3)
    acc[::7, 3] = 2.0
    acc[::11, 3] = 0.0
    i0, i1 = 150, 4800
    a, _ = oracle.fmm_sph(pos, vel, acc, p=10, max_level=2, i0=i0, i1=i1, rotational=True)
    ref = oracle.direct_ref(pos, vel, acc, i0=i0, i1=i1, j0=i0, j1=i1, sym="rotational", nsec=4)
    tgt = acc[i0:i1, 3] == 1.0
    assert np.all(a[~tgt] == 0.0)
    assert oracle.rel_l2(a[tgt], ref[tgt])[0] < 1e-5


# ---- translational FMM (ternary tree over one periodic cell) ----

LAT27 = np.array([(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)], np.int32)


def _cell(n, cell, seed, clump=0):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-cell / 2, cell / 2, (n, 3))
    if clump:
        pos[:clump, :3] = cell / 2 - np.abs(rng.normal(0, cell / 40, (clump, 3)))
    pos[:, 3] = np.where(rng.random(n) < 0.1, -1.0, 1.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = rng.integers(0, 2, n)
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    return pos, vel, acc


@pytest.mark.parametrize("level", [1, 2])
def test_translational_fmm_converges_to_27_image_direct(level):
    """Every target of the translational FMM sees the sources of the 27 cells
    at shifts {-1, 0, 1}^3 (near field through the wrapped P2P list, far field
    through the wrapped M2L lists of every level): the error against the FP64
    direct sum over those images falls with the order.  A wrong ternary shift
    (rho, 3^-j), rotation index or list wrap stalls it."""
    L = 2e-6
    pos, vel, acc = _cell(1500, L, 5, clump=300)
    acc[::7, 3] = 2.0
    ref = oracle.direct_ref(pos, vel, acc, sym="translational", lattice=LAT27, domain=(L, L, L))
    tgt = acc[:, 3] == 1.0
    errs = []
    for p in (4, 6, 8, 10):
        out, info = oracle.fmm_sph_trans(pos, vel, acc, L, p=p, max_level=level, cr_trig=True)
        errs.append(oracle.rel_l2(out[tgt], ref[tgt])[0])
    assert info["max_level"] == level and info["leaves"] <= 27 ** level
    assert all(b < a / 3 for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 5e-4, errs


def test_translational_near_field_and_reference_p2p():
    """near=False is the reference as it runs (its translational P2P passes the
    thread's box range as the ion range and adds nothing); the difference to
    near=True is the softened pair sum over the 27 wrapped neighbour leaves,
    restated here by brute force from the leaf indices of mortonTranslations."""
    L = np.float32(2e-6)
    pos, vel, acc = _cell(400, float(L), 6, clump=80)
    acc[::5, 3] = 2.0
    acc[::13, 3] = 0.0
    full, info = oracle.fmm_sph_trans(pos, vel, acc, float(L), p=6, max_level=2)
    far, _ = oracle.fmm_sph_trans(pos, vel, acc, float(L), p=6, max_level=2, near=False)
    side = 9
    bmin = np.float32(-L / np.float32(2))
    bs = np.float32(L / np.float32(side))
    cellidx = np.minimum(((pos[:, :3] - bmin) / bs).astype(np.int64), side - 1)
    src = (acc[:, 3] == 1.0) | (acc[:, 3] == 2.0)
    kc = oracle.k_const()
    near = np.zeros((pos.shape[0], 3))
    for i in np.nonzero(acc[:, 3] == 1.0)[0]:
        ac = kc * float(pos[i, 3]) / oracle.species_mass(int(vel[i, 3]))
        for j in np.nonzero(src)[0]:
            d = cellidx[j] - cellidx[i]
            d = (d + 1) % side - 1  # periodic leaf distance in [-1, side - 2]
            if np.any(np.abs(d) > 1):
                continue
            # image shift of the source leaf as the wrap of cellidx[i] + d finds it
            s = np.floor_divide(cellidx[i] + d, side)
            xi = pos[i, :3] - s.astype(np.float32) * L  # float, as direct_translational
            dd = (xi - pos[j, :3]).astype(np.float64)
            inv = 1.0 / math.sqrt(float(dd @ dd) + 6.4e-19)
            near[i] += ac * float(pos[j, 3]) * inv ** 3 * dd
    tgt = acc[:, 3] == 1.0
    assert np.all(far[~tgt] == 0.0) and np.all(full[~tgt] == 0.0)
    np.testing.assert_allclose(full[tgt] - far[tgt], near[tgt], rtol=1e-9, atol=1e-9 * np.abs(near).max())
    assert oracle.rel_l2(far[tgt], full[tgt])[0] > 0.1


def test_translational_level_and_unsupported():
    L = 2e-6
    pos, vel, acc = _cell(1500, L, 7)
    _, info = oracle.fmm_sph_trans(pos, vel, acc, L, p=2)
    assert info["max_level"] == 2 and (info["nd0"], info["nd1"], info["nd2"]) == (1, 1, 1)
    _, info = oracle.fmm_sph_trans(pos[:150], vel[:150], acc[:150], L, p=2, max_level=1)
    assert info["leaves"] <= 27 and info["boxes"] == info["leaves"]
    with pytest.raises(ValueError):  # 150 sources: level 0, the supercell-only path
        oracle.fmm_sph_trans(pos[:150], vel[:150], acc[:150], L, p=2)
    with pytest.raises(ValueError):
        oracle.fmm_sph_trans(pos, vel, acc, L, p=2, max_level=7)
    out = pos.copy()
    out[3, 1] = np.float32(L / 2)  # on the +y face: outside [-L/2, L/2)
    with pytest.raises(ValueError):
        oracle.fmm_sph_trans(out, vel, acc, L, p=2)
