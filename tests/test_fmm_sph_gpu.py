"""Parity of the GPU spherical-harmonic FMM (the reference FMM's own
operators: rotation, z shift, rotation back; shp_fmm_sph_host) with the
oracle's FP64 restatement of them (oracle/fmm_sph.c) on the same float AoS
inputs, through the C ABI.

Both sides are FP64 and build the same tree (the reference's domain grid,
level choice and float leaf placement), so the tree summaries must be equal
and the accelerations agree to rounding: the GPU applies each translation as
one dense operator on the FP64 matrix cores, where the oracle rotates, shifts
and rotates back per box pair.

cart2sph_exact takes float acosf / atanf of its arguments, so the reference's
result depends on its libm in the last float bit of each angle.  The GPU
rounds the double functions to float (correctly rounded); this host's glibc
acosf differs from that on ~7% of arguments.  Bars: against the oracle with
correctly rounded angles, rel-L2 <= 1e-12 and max per-ion relative difference
<= 1e-10; against the oracle with this host's libm (the reference's own
arithmetic on x86-64 glibc), rel-L2 <= 1e-7 (the libm boundary; 5e-7 for the
translational FMM, whose far field carries most of each target's force)."""
import numpy as np
import pytest

import oracle
from solveit_hip import synth

pytestmark = pytest.mark.gpu

# the 27 lattice points {-1, 0, 1}^3 of the reference's translational FMM
LAT27 = np.array([(x, y, z) for x in (-1, 0, 1) for y in (-1, 0, 1) for z in (-1, 0, 1)], np.int32)


@pytest.fixture(scope="module")
def eng():
    import solveit_hip
    import torch
    assert torch.cuda.is_available()
    e = solveit_hip.Engine(0)
    yield e
    e.close()


def _cube(n, side, seed, zscale=1.0):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, side, (n, 3))
    pos[:, 2] *= zscale
    pos[:, 3] = np.where(rng.random(n) < 0.05, -1.0, 1.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = rng.integers(0, 2, n)
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    return pos, vel, acc


def _check(eng, pos, vel, acc, p, level=-2, i0=0, i1=None, label="", rotational=False, cell=None):
    import solveit_hip
    i1 = pos.shape[0] if i1 is None else i1
    print(f"case {label}: n={pos.shape[0]} p={p} level={level} range=[{i0},{i1}) rotational={rotational} "
          f"cell={cell}", flush=True)
    sym = solveit_hip.make_symmetry("rotational", nsec=4) if rotational else None
    if cell is not None:
        sym = solveit_hip.make_symmetry("translational", lattice=LAT27, domain=(cell, cell, cell))
    got = eng.fmm_sph_host(pos, vel, acc, i0=i0, i1=i1, order=p, max_level=level, symmetry=sym)
    st = eng.fmm_sph_stats()
    if cell is not None:
        ref, info = oracle.fmm_sph_trans(pos, vel, acc, cell, p=p, max_level=level, i0=i0, i1=i1, cr_trig=True)
        ref_libm, _ = oracle.fmm_sph_trans(pos, vel, acc, cell, p=p, max_level=level, i0=i0, i1=i1)
    else:
        ref, info = oracle.fmm_sph(pos, vel, acc, p=p, max_level=level, i0=i0, i1=i1, cr_trig=True,
                                   rotational=rotational)
        ref_libm, _ = oracle.fmm_sph(pos, vel, acc, p=p, max_level=level, i0=i0, i1=i1, rotational=rotational)
    for k in ("max_level", "k", "nd0", "nd1", "nd2", "leaves", "m2l_pairs", "p2p_pairs", "sources",
              "boxes", "p2p_ion_pairs"):
        assert st[k] == info[k], (k, st[k], info[k])
    tgt = acc[i0:i1, 3] == 1.0
    assert np.all(got[~tgt] == 0.0)
    l2 = oracle.rel_l2(got[tgt], ref[tgt])[0]
    mx = np.max(np.linalg.norm(got[tgt] - ref[tgt], axis=1) / np.linalg.norm(ref[tgt], axis=1))
    l2m = oracle.rel_l2(got[tgt], ref_libm[tgt])[0]
    print(f"  rel L2 {l2:.2e}, max {mx:.2e} (host libm: rel L2 {l2m:.2e}), stats {st}", flush=True)
    assert l2 <= 1e-12 and mx <= 1e-10, (l2, mx)
    # the libm boundary: the translational far field (26 of the 27 cells reach a
    # target through M2L only) carries more of the float-angle rounding
    assert l2m <= (5e-7 if cell is not None else 1e-7), l2m
    return got, ref


@pytest.mark.parametrize("p", [1, 4, 8, 10])
def test_single_domain_cube(eng, p):
    """One domain (the reference's tree exactly), level from setOptimumLevel."""
    pos, vel, acc = _cube(6000, 1e-5, 4)
    _check(eng, pos, vel, acc, p, label="cube 10um")


def test_single_domain_forced_level(eng):
    pos, vel, acc = _cube(20000, 1e-5, 6)
    _check(eng, pos, vel, acc, 10, level=4, label="cube 10um L4")


def test_multi_domain_slab(eng):
    """A 4:4:1 slab: a 4 x 4 x 1 domain grid, octree levels above it."""
    pos, vel, acc = _cube(6000, 1e-5, 7, zscale=0.25)
    _check(eng, pos, vel, acc, 10, level=2, label="slab")


def test_multi_domain_column(eng):
    """A column of six domains along z: the float domain corners bmin + nd *
    rootBoxSize, which the reference rounds product first; a fused
    multiply-add gives other corners for nd = 3 and 5 on this input."""
    pos, vel, acc = _cube(6000, 1e-5, 7)
    pos[:, 2] *= 6.0
    _check(eng, pos, vel, acc, 8, level=3, label="column, 6 domains")


def test_cone_plume_and_accuracy(eng):
    """The synthetic cone (a chain of domains along z): parity with the oracle,
    and the reference expansion's own accuracy against the direct sum."""
    d = synth.cone(16384, seed=11)
    pos, vel, acc = synth.to_reference_aos(d)
    got, ref = _check(eng, pos, vel, acc, 10, label="cone 16k")
    direct = oracle.direct_ref(pos, vel, acc)
    err = oracle.rel_l2(got, direct)[0]
    print(f"  cone 16k p=10 vs direct: {err:.2e}", flush=True)
    assert err < 1e-3


def test_status_rules_and_range(eng):
    pos, vel, acc = _cube(3000, 1e-5, 8)
    acc[::7, 3] = 2.0
    acc[::11, 3] = 0.0
    _check(eng, pos, vel, acc, 8, level=2, i0=100, i1=2900, label="statuses")


def test_dense_cube_softening_floor(eng):
    """100 nm cube: the leaf boxes are 16x the softening length, where the
    reference's plain-1/r expansion stops converging; the GPU reproduces the
    oracle there too."""
    pos, vel, acc = _cube(8000, 1e-7, 5)
    _check(eng, pos, vel, acc, 10, label="dense cube")


def test_unsupported_cases(eng):
    import solveit_hip
    pos, vel, acc = _cube(100, 1e-5, 9)
    with pytest.raises(solveit_hip.ShpError):  # maxLevel 0 from 100 sources
        eng.fmm_sph_host(pos, vel, acc, order=4)
    with pytest.raises(solveit_hip.ShpError):
        eng.fmm_sph_host(pos, vel, acc, order=11, max_level=2)


def test_engine_fmm_against_reference_operators(eng):
    """The engine's own FMM (Cartesian expansion of the softened kernel) and the
    reference's operators at the same order on the same float inputs agree to
    the reference expansion's own error against the direct sum (its plain-1/r
    expansion stops near 2e-5 on this cone), and the engine is the closer of
    the two to the direct sum."""
    d = synth.cone(16384, seed=11)
    pos, vel, acc = synth.to_reference_aos(d)
    ref_ops = eng.fmm_sph_host(pos, vel, acc, order=10)
    mine = eng.fmm_host(pos, vel, acc, order=10, ring=2)
    direct = oracle.direct_ref(pos, vel, acc)
    e_ref = oracle.rel_l2(ref_ops, direct)[0]
    e_mine = oracle.rel_l2(mine, direct)[0]
    between = oracle.rel_l2(mine, ref_ops)[0]
    print(f"  p=10 cone 16k: reference operators {e_ref:.2e}, engine {e_mine:.2e}, between {between:.2e}",
          flush=True)
    assert e_mine < e_ref
    assert between <= 2.0 * (e_ref + e_mine)


# ---- rotational symmetry: the reference's four quadrant images ----

def _quadrant_column(n, side, zlen, seed):
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


@pytest.mark.parametrize("p", [4, 10])
def test_rotational_single_domain(eng, p):
    pos, vel, acc = _quadrant_column(6000, 1e-5, 1.5e-5, 1)
    _check(eng, pos, vel, acc, p, label="quadrant column, one domain", rotational=True)


def test_rotational_domain_column(eng):
    """Six domains along z (K = 3): the four-image M2M into each domain, plain
    octree levels above it, the quadrant map below."""
    pos, vel, acc = _quadrant_column(6000, 1e-5, 1.2e-4, 1)
    _check(eng, pos, vel, acc, 8, level=3, label="quadrant column, 6 domains", rotational=True)


def test_rotational_folded_cone(eng):
    """The cone folded into the first quadrant: parity with the oracle, and
    the reference expansion's accuracy against the direct sum over the four
    quadrant images."""
    d = synth.fold_sectors(synth.cone(16384, seed=11), 4)
    pos, vel, acc = synth.to_reference_aos(d)
    acc[::13, 3] = 2.0
    got, _ = _check(eng, pos, vel, acc, 10, label="folded cone 16k", rotational=True)
    direct = oracle.direct_ref(pos, vel, acc, sym="rotational", nsec=4)
    tgt = acc[:, 3] == 1.0
    err = oracle.rel_l2(got[tgt], direct[tgt])[0]
    print(f"  folded cone 16k p=10 vs rotational direct: {err:.2e}", flush=True)
    assert err < 1e-3


def test_rotational_status_rules_and_range(eng):
    """Rotational with status-2 sources, status-0 ions and a sub-range: the
    images are those of the sources in the range only."""
    pos, vel, acc = _quadrant_column(5000, 1e-5, 1.5e-5, 3)
    acc[::7, 3] = 2.0
    acc[::11, 3] = 0.0
    _check(eng, pos, vel, acc, 8, level=2, i0=150, i1=4800, label="rotational statuses", rotational=True)


def test_rotational_unsupported(eng):
    import solveit_hip
    pos, vel, acc = _quadrant_column(3000, 1e-5, 1e-5, 2)  # a 2 x 2 x 1 rotational grid
    with pytest.raises(solveit_hip.ShpError):
        eng.fmm_sph_host(pos, vel, acc, order=4, symmetry=solveit_hip.make_symmetry("rotational", nsec=4))
    pos, vel, acc = _quadrant_column(6000, 1e-5, 1.5e-5, 1)
    with pytest.raises(solveit_hip.ShpError):
        eng.fmm_sph_host(pos, vel, acc, order=4, symmetry=solveit_hip.make_symmetry("rotational", nsec=6))


def _cell(n, cell, seed, clump=0):
    """Ions uniform in the periodic cell [-cell/2, cell/2)^3; `clump` of them
    packed against the +x +y +z corner, so their near and far fields reach
    through the periodic wrap."""
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-cell / 2, cell / 2, (n, 3))
    if clump:
        pos[:clump, :3] = cell / 2 - np.abs(rng.normal(0, cell / 40, (clump, 3)))
        pos[:clump, :3] = np.minimum(pos[:clump, :3], np.nextafter(np.float32(cell / 2), np.float32(0)))
    pos[:, 3] = np.where(rng.random(n) < 0.05, -1.0, 1.0)
    vel = np.zeros((n, 4), np.float32)
    vel[:, 3] = rng.integers(0, 2, n)
    acc = np.zeros((n, 4), np.float32)
    acc[:, 3] = 1.0
    return pos, vel, acc


@pytest.mark.parametrize("p,level", [(4, -2), (10, -2), (6, 1), (8, 3)])
def test_translational_cell(eng, p, level):
    """The reference's translational FMM: ternary tree of one periodic cell,
    levels from 1, lists wrapped with the image shift."""
    pos, vel, acc = _cell(4000, 2e-6, 21, clump=600)
    _check(eng, pos, vel, acc, p, level=level, label="periodic cell", cell=2e-6)


def test_translational_against_direct_images(eng):
    """Accuracy of the ternary expansion against the FP64 direct sum over the
    27 images (what every target of the translational FMM sees)."""
    pos, vel, acc = _cell(6000, 2e-6, 22, clump=1000)
    acc[::9, 3] = 2.0
    got, _ = _check(eng, pos, vel, acc, 10, label="cell vs direct", cell=2e-6)
    direct = oracle.direct_ref(pos, vel, acc, sym="translational", lattice=LAT27, domain=(2e-6, 2e-6, 2e-6))
    tgt = acc[:, 3] == 1.0
    err = oracle.rel_l2(got[tgt], direct[tgt])[0]
    print(f"  cell 6k p=10 vs 27-image direct: {err:.2e}", flush=True)
    assert err < 1e-3


def test_translational_status_rules_and_range(eng):
    pos, vel, acc = _cell(5000, 2e-6, 23, clump=500)
    acc[::7, 3] = 2.0
    acc[::11, 3] = 0.0
    _check(eng, pos, vel, acc, 8, level=2, i0=120, i1=4700, label="translational statuses", cell=2e-6)


def test_translational_unsupported(eng):
    import solveit_hip
    pos, vel, acc = _cell(3000, 2e-6, 24)
    sym = solveit_hip.make_symmetry("translational", lattice=LAT27, domain=(2e-6, 2e-6, 2e-6))
    for lvl in (0, 7):  # the supercell-only path; the reference's int box index overflows at 27^7
        with pytest.raises(solveit_hip.ShpError):
            eng.fmm_sph_host(pos, vel, acc, order=4, max_level=lvl, symmetry=sym)
    with pytest.raises(solveit_hip.ShpError):  # not a cube
        eng.fmm_sph_host(pos, vel, acc, order=4, symmetry=solveit_hip.make_symmetry(
            "translational", lattice=LAT27, domain=(2e-6, 2e-6, 3e-6)))
    with pytest.raises(solveit_hip.ShpError):  # another image set
        eng.fmm_sph_host(pos, vel, acc, order=4, symmetry=solveit_hip.make_symmetry(
            "translational", lattice=LAT27[:9], domain=(2e-6, 2e-6, 2e-6)))
    out = pos.copy()
    out[5, 0] = 1e-6  # on the +x face: outside [-cell/2, cell/2) (the reference's outlier passes)
    with pytest.raises(solveit_hip.ShpError):
        eng.fmm_sph_host(out, vel, acc, order=4, symmetry=sym)
