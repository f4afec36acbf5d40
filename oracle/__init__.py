"""Oracle: CPU restatement of the reference's force-engine hot path.

TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg may import this package, and only as the checker (or as the
timed CPU baseline).  The product library never links or calls it.

The arithmetic lives in C (oracle/*.c, built by oracle/Makefile into
oracle/liboracle.so with -ffp-contract=off); this module is its ctypes front end
plus the pure-Python restatements of the reference's small index algorithms.

Pinning (what each function is checked against, see tests/test_oracle_*.py):
  orc_k_const            bit pattern 0x1.1c8a2aa5ab9edp-3 of K_CONST (x86-64 g++).
  direct (ref, f64)      two- and three-body closed forms with softening,
                         momentum identity sum m_i a_i = 0, long-double brute
                         force, float replay of the reference's accumulation.
  rotational sectors     equality with the plain sum over explicitly rotated
                         source copies; 4-fold integer cos/sin table.
  translational images   equality with the plain sum over explicitly shifted
                         source copies; image set of reference_lattice() equals
                         the full (2n+1)^d block, each point once.
  rel_l2                 hand cases (identical, scaled, zero-reference rows).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    ip = ctypes.POINTER(ctypes.c_int)
    ci, cd = ctypes.c_int, ctypes.c_double
    lib.orc_k_const.restype = cd
    lib.orc_softening.restype = cd
    lib.orc_e_over_amu.restype = cd
    lib.orc_species_mass.restype = cd
    lib.orc_species_mass.argtypes = [ci]
    lib.orc_quadrant_cos.argtypes = [ci]
    lib.orc_quadrant_sin.argtypes = [ci]
    lib.orc_direct_ref.argtypes = [fp, fp, fp, ci, ci, ci, ci, ci, ci, ci, ip, fp, dp]
    lib.orc_direct_ref_accumulate.argtypes = [fp, fp, fp, ci, ci, ci, ci, ci, ci, ci, ip, fp]
    lib.orc_direct_f64.argtypes = [ci, dp, dp, dp, dp, ip, ip, ci, ci, ci, ci, ci, ip, dp, dp]
    lib.orc_direct_f64_ld.argtypes = [ci, dp, dp, dp, dp, ip, ip, dp]
    return lib


lib = _load()

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def k_const():
    return lib.orc_k_const()


def softening():
    return lib.orc_softening()


def species_mass(s):
    return lib.orc_species_mass(int(s))


_SYM = {"none": 0, "rotational": 1, "translational": 2}


def direct_ref(pos4, vel4, acc4, i0=0, i1=None, j0=0, j1=None, sym="none", nsec=4,
               lattice=None, domain=(0.0, 0.0, 0.0)):
    """Reference-contract direct sum (pre-rounding double), rows i0..i1."""
    pos4, vel4, acc4 = _f32(pos4), _f32(vel4), _f32(acc4)
    n = pos4.shape[0]
    i1 = n if i1 is None else i1
    j1 = n if j1 is None else j1
    lat = _i32(np.zeros((1, 3)) if lattice is None else lattice)
    dom = _f32(domain)
    out = np.zeros((i1 - i0, 3))
    lib.orc_direct_ref(pos4.ctypes.data_as(_fp), vel4.ctypes.data_as(_fp),
                       acc4.ctypes.data_as(_fp), i0, i1, j0, j1, _SYM[sym], nsec,
                       lat.shape[0], lat.ctypes.data_as(_ip), dom.ctypes.data_as(_fp),
                       out.ctypes.data_as(_dp))
    return out


def direct_ref_accumulate(pos4, vel4, acc4, i0=0, i1=None, j0=0, j1=None, sym="none", nsec=4,
                          lattice=None, domain=(0.0, 0.0, 0.0)):
    """Replays the reference's float accumulation; returns a new acc4 array."""
    pos4, vel4 = _f32(pos4), _f32(vel4)
    acc = _f32(acc4).copy()
    n = pos4.shape[0]
    i1 = n if i1 is None else i1
    j1 = n if j1 is None else j1
    lat = _i32(np.zeros((1, 3)) if lattice is None else lattice)
    dom = _f32(domain)
    lib.orc_direct_ref_accumulate(pos4.ctypes.data_as(_fp), vel4.ctypes.data_as(_fp),
                                  acc.ctypes.data_as(_fp), i0, i1, j0, j1, _SYM[sym], nsec,
                                  lat.shape[0], lat.ctypes.data_as(_ip), dom.ctypes.data_as(_fp))
    return acc


def direct_f64(d, i0=0, i1=None, sym="none", nsec=4, lattice=None, domain=(0.0, 0.0, 0.0)):
    """FP64-state direct sum on a synth dict (x, y, z, q, status, species)."""
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    n = x.shape[0]
    i1 = n if i1 is None else i1
    lat = _i32(np.zeros((1, 3)) if lattice is None else lattice)
    dom = _f64(domain)
    out = np.zeros((i1 - i0, 3))
    lib.orc_direct_f64(n, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp), z.ctypes.data_as(_dp),
                       q.ctypes.data_as(_dp), st.ctypes.data_as(_ip), sp.ctypes.data_as(_ip),
                       i0, i1, _SYM[sym], nsec, lat.shape[0], lat.ctypes.data_as(_ip),
                       dom.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
    return out


def direct_f64_ld(d):
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    n = x.shape[0]
    out = np.zeros((n, 3))
    lib.orc_direct_f64_ld(n, x.ctypes.data_as(_dp), y.ctypes.data_as(_dp),
                          z.ctypes.data_as(_dp), q.ctypes.data_as(_dp),
                          st.ctypes.data_as(_ip), sp.ctypes.data_as(_ip),
                          out.ctypes.data_as(_dp))
    return out


def _oddeven(n):
    return -1 if (n & 1) else 1


def reference_grid(box_length, axes):
    """Restatement of Grid(boxLength, symmetryAxes), lattice_point.h:65-131:
    non-negative lattice points on the symmetry axes, sorted by decreasing norm."""
    assert box_length % 2 == 1
    n = box_length // 2
    s = int(sum(axes))
    num = {1: n + 1, 2: n * (n + 1) + 1, 3: n * (n * n + n + 1) + 1}.get(s, 1)
    lat = []
    t = list(axes)
    for _ in range(s):
        i = t[0]
        while True:
            j = t[1]
            while True:
                k = t[2]
                while True:
                    lat.append((i, j, k))
                    k += 1
                    if not (t[2] != 0 and k <= n):
                        break
                j += 1
                if not (t[1] != 0 and j <= n):
                    break
            i += 1
            if not (t[0] != 0 and i <= n):
                break
        if t[0] == 0:
            if t[1] == 0:
                t[2] = 0
            else:
                t[1] = 0
        else:
            t[0] = 0
    lat.append((0, 0, 0))
    assert len(lat) == num
    # std::sort with std::greater on the norm (ties keep no particular order)
    lat.sort(key=lambda p: -(p[0] ** 2 + p[1] ** 2 + p[2] ** 2))
    return lat, s


def reference_lattice(box_length, axes):
    """Image offsets in the order DirectForceCalculation::directKernelTranslationCalculation
    visits them (direct_force_method_translational.cpp:28-309)."""
    lat, s = reference_grid(box_length, axes)
    ax = axes
    out = []
    for (px, py, pz) in lat:
        nz = (px != 0) + (py != 0) + (pz != 0)
        rx, ry, rz = range(1 + (px != 0)), range(1 + (py != 0)), range(1 + (pz != 0))
        if s == 3 and nz == 3:
            for i in range(2):
                for j in range(2):
                    for k in range(2):
                        out.append((_oddeven(i) * px, _oddeven(j) * py, _oddeven(k) * pz))
        elif s == 3 and nz == 2:
            for jj in range(3):
                for i in rx:
                    for j in ry:
                        for k in rz:
                            oi, oj, ok = _oddeven(i), _oddeven(j), _oddeven(k)
                            out.append((
                                (jj != 0) * oi * px + (px == 0) * ((jj == 1) * oj * py + (jj == 2) * ok * pz),
                                (jj != 1) * oj * py + (py == 0) * ((jj == 0) * oi * px + (jj == 2) * ok * pz),
                                (jj != 2) * ok * pz + (pz == 0) * ((jj == 0) * oi * px + (jj == 1) * oj * py)))
        elif s == 3 and nz == 1:
            for jj in range(3):
                for i in rx:
                    for j in ry:
                        for k in rz:
                            oi, oj, ok = _oddeven(i), _oddeven(j), _oddeven(k)
                            out.append((
                                (jj == 0) * oi * px + (jj == 1) * oj * py + (jj == 2) * ok * pz,
                                (jj == 0) * oj * py + (jj == 1) * ok * pz + (jj == 2) * oi * px,
                                (jj == 0) * ok * pz + (jj == 1) * oi * px + (jj == 2) * oj * py))
        elif s == 2 and nz == 2:
            for i in rx:
                for j in ry:
                    for k in rz:
                        out.append((_oddeven(i) * px, _oddeven(j) * py, _oddeven(k) * pz))
        elif s == 2 and nz == 1:
            for jj in range(2):
                for i in rx:
                    for j in ry:
                        for k in rz:
                            oi, oj, ok = _oddeven(i), _oddeven(j), _oddeven(k)
                            out.append((
                                ax[0] * ((jj == 0) * oi * px + (jj == 1) * (oj * py + ok * pz)),
                                ax[1] * ((jj == 0) * oj * py + (jj == 1) * (oi * px + ok * pz)),
                                ax[2] * ((jj == 0) * ok * pz + (jj == 1) * (oi * px + oj * py))))
        elif s == 1 and nz == 1:
            for i in rx:
                for j in ry:
                    for k in rz:
                        out.append((_oddeven(i) * px, _oddeven(j) * py, _oddeven(k) * pz))
        else:
            out.append((0, 0, 0))
    return np.array(out, dtype=np.int32)


def rel_l2(a, ref, mask=None):
    """Relative L2 of Integrator::compareAcceleration (integration_schemes.cpp:411-455):
    sqrt(mean_i |a_i - ref_i|^2 / |ref_i|^2) over status-1 rows, rows with both
    terms zero counting 0; also returns the max per-row ratio (sqrt) and its index.
    Evaluated in float64."""
    a = np.asarray(a, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    if mask is not None:
        a, ref = a[mask], ref[mask]
    diff = np.sum((a - ref) ** 2, axis=1)
    norm = np.sum(ref ** 2, axis=1)
    ratio = np.where((norm == 0) & (diff == 0), 0.0, diff / np.where(norm == 0, 1.0, norm))
    ratio = np.where((norm == 0) & (diff != 0), np.inf, ratio)
    if ratio.size == 0:
        return 0.0, 0.0, -1
    k = int(np.argmax(ratio))
    return float(np.sqrt(ratio.mean())), float(np.sqrt(ratio[k])), k


# ---------------------------------------------------------------------------
# Background field (oracle/pwl.c) and leapfrog (oracle/leapfrog.c)
# ---------------------------------------------------------------------------
class _LfParams(ctypes.Structure):
    _fields_ = [
        ("sym", ctypes.c_int), ("nsec", ctypes.c_int),
        ("dom", ctypes.c_double * 3),
        ("bc", ctypes.c_double * 4),
        ("bc_time", ctypes.c_double),
        ("use_efield", ctypes.c_int),
        ("xy", ctypes.c_void_p), ("er", ctypes.c_void_p), ("ez", ctypes.c_void_p),
        ("tri", ctypes.c_void_p), ("nbr", ctypes.c_void_p),
        ("ntri", ctypes.c_int),
        ("r0", ctypes.c_double), ("E0", ctypes.c_double),
        ("nimg", ctypes.c_int), ("lattice", ctypes.c_void_p),
    ]


_cd, _ci, _vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
lib.orc_tri_search.restype = _ci
lib.orc_tri_search.argtypes = [_vp, _ci, _vp, _vp, _vp, _vp, _vp, _vp]
lib.orc_efield_ref.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                               ctypes.c_float, _vp, _ci, _vp, _vp, _vp, _vp, _vp, _vp]
lib.orc_background_accel_ref.argtypes = [_vp, _vp, _vp, _ci, _ci, ctypes.c_float, ctypes.c_float,
                                         _vp, _ci, _vp, _vp, _vp, _vp, _vp, _vp]
lib.orc_efield_f64_batch.argtypes = [_ci, _vp, _vp, _vp, _cd, _cd, _vp, _ci, _vp, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp]
lib.orc_leapfrog_run_f64.argtypes = [ctypes.POINTER(_LfParams), _ci, _vp, _vp, _vp, _vp, _vp, _vp,
                                     _vp, _vp, _vp, _vp, _cd, _cd, _ci, _vp]


def _mesh_arrays(mesh):
    return {"xy": _f64(mesh["xy"]), "er": _f64(mesh["er"]), "ez": _f64(mesh["ez"]),
            "tri": _i32(mesh["tri"]), "nbr": _i32(mesh["nbr"])}


def efield_f64(mesh, x, y, z, seeds=None):
    """FP64 field at points (walk from `seeds` or the default start).
    Returns (E[n,3], edge[n], tri[n])."""
    m = _mesh_arrays(mesh)
    x, y, z = _f64(x), _f64(y), _f64(z)
    n = x.shape[0]
    out = np.zeros((n, 3))
    edges = np.zeros(n, np.int32)
    tris = np.zeros(n, np.int32)
    sd = None if seeds is None else _i32(seeds)
    lib.orc_efield_f64_batch(n, x.ctypes.data, y.ctypes.data, z.ctypes.data, float(mesh["r0"]),
                             float(mesh["E0"]), m["xy"].ctypes.data, m["tri"].shape[0],
                             m["tri"].ctypes.data, m["nbr"].ctypes.data, m["er"].ctypes.data,
                             m["ez"].ctypes.data, None if sd is None else sd.ctypes.data,
                             out.ctypes.data, edges.ctypes.data, tris.ctypes.data)
    return out, edges, tris


def tri_search(mesh, p, start=-1):
    """One reference walk from `start`; returns (edge, tri, (alpha, beta, gamma), steps)."""
    m = _mesh_arrays(mesh)
    pp = _f64(p)
    t = ctypes.c_int(start)
    bc = np.zeros(3)
    steps = ctypes.c_int(0)
    e = lib.orc_tri_search(m["xy"].ctypes.data, m["tri"].shape[0], m["tri"].ctypes.data,
                           m["nbr"].ctypes.data, pp.ctypes.data, ctypes.addressof(t),
                           bc.ctypes.data, ctypes.addressof(steps))
    return e, t.value, tuple(bc), steps.value


def background_accel_ref(mesh, pos4, vel4, acc4, start=-1):
    """computeBackgroundAcceleration on the reference's float arrays, in index
    order with the warm start chained; returns (acc4', efield[n,3] float32)."""
    m = _mesh_arrays(mesh)
    pos4, vel4 = _f32(pos4), _f32(vel4)
    acc = _f32(acc4).copy()
    n = pos4.shape[0]
    ef = np.zeros((n, 3), np.float32)
    t = ctypes.c_int(start)
    lib.orc_background_accel_ref(pos4.ctypes.data, vel4.ctypes.data, acc.ctypes.data, 0, n,
                                 float(np.float32(mesh["r0"])), float(np.float32(mesh["E0"])),
                                 m["xy"].ctypes.data, m["tri"].shape[0], m["tri"].ctypes.data,
                                 m["nbr"].ctypes.data, m["er"].ctypes.data, m["ez"].ctypes.data,
                                 ctypes.addressof(t), ef.ctypes.data)
    return acc, ef


def _lf_params(mesh, sym, nsec, lattice, domain, bc_space, bc_time):
    P = _LfParams()
    P.sym = _SYM[sym]
    P.nsec = nsec
    for c in range(3):
        P.dom[c] = float(domain[c])
    for c in range(4):
        P.bc[c] = float(bc_space[c])
    P.bc_time = float(bc_time)
    keep = []
    if mesh is not None:
        m = _mesh_arrays(mesh)
        keep.append(m)
        P.use_efield = 1
        P.xy, P.er, P.ez = m["xy"].ctypes.data, m["er"].ctypes.data, m["ez"].ctypes.data
        P.tri, P.nbr = m["tri"].ctypes.data, m["nbr"].ctypes.data
        P.ntri = m["tri"].shape[0]
        P.r0, P.E0 = float(mesh["r0"]), float(mesh["E0"])
    lat = _i32(np.zeros((1, 3)) if lattice is None else lattice)
    keep.append(lat)
    P.nimg = lat.shape[0]
    P.lattice = lat.ctypes.data
    return P, keep


_STATE_KEYS = ("x", "y", "z", "vx", "vy", "vz", "q", "status", "species", "t_inj")


def _state_copy(d):
    return {k: (_f64(d[k]).copy() if k not in ("status", "species") else _i32(d[k]).copy())
            for k in _STATE_KEYS}


def leapfrog_run_f64(d, t0, dt, nsteps, mesh=None, sym="none", nsec=4, lattice=None,
                     domain=(0.0, 0.0, 0.0), bc_space=(0, 0, 0, 0), bc_time=0.0):
    """nsteps of DKD (FP64 direct force) on a copy of the synth dict; returns
    (state dict, last Coulomb acceleration[n,3])."""
    s = _state_copy(d)
    n = s["x"].shape[0]
    P, keep = _lf_params(mesh, sym, nsec, lattice, domain, bc_space, bc_time)
    acc = np.zeros((n, 3))
    lib.orc_leapfrog_run_f64(ctypes.byref(P), n, *[s[k].ctypes.data for k in _STATE_KEYS],
                             float(t0), float(dt), int(nsteps), acc.ctypes.data)
    return s, acc


class SubstepLevel(ctypes.Structure):
    """One level of LeapfrogDKDSubsteps' boundary-condition table: bc_type
    0 none, 1 space, 2 time, 3 space and time; bc_space |x|, |y|, |z|, rho
    limits; bc_time limit on the time since injection (0: no limit)."""
    _fields_ = [("bc_type", ctypes.c_int), ("bc_space", ctypes.c_double * 4), ("bc_time", ctypes.c_double)]


def substep_table(levels):
    """levels: list of (bc_type, (bx, by, bz, brho), bt) for levels 0..M."""
    arr = (SubstepLevel * len(levels))()
    for k, (typ, sp, tm) in enumerate(levels):
        arr[k].bc_type = int(typ)
        for c in range(4):
            arr[k].bc_space[c] = float(sp[c])
        arr[k].bc_time = float(tm)
    return arr


lib.orc_substep_level_of.restype = _ci
lib.orc_substep_level_of.argtypes = [_vp, _ci, _ci, _ci, _cd, _cd, _cd, _cd, _cd]
lib.orc_leapfrog_substeps_f64.argtypes = [ctypes.POINTER(_LfParams), _vp, _ci, _ci] + [_vp] * 11 + \
    [_cd, _cd, _ci, _vp]


def substep_level_of(levels, level, cur, x, y, z, tnow, tinj):
    """checkConditions of LeapfrogDKDSubsteps for one ion: its new level."""
    T = substep_table(levels)
    return lib.orc_substep_level_of(ctypes.addressof(T), len(levels) - 1, int(level), int(cur), float(x),
                                    float(y), float(z), float(tnow), float(tinj))


def leapfrog_substeps_f64(d, t0, dt, nsteps, levels, level0=None, mesh=None, sym="none", nsec=4,
                          lattice=None, domain=(0.0, 0.0, 0.0), bc_space=(0, 0, 0, 0), bc_time=0.0):
    """nsteps of LeapfrogDKDSubsteps (FP64 direct force) on a copy of the synth
    dict with the level table `levels` (levels 0..M); level0: the ions' levels
    at the start (default M).  Returns (state dict, levels, last force)."""
    s = _state_copy(d)
    n = s["x"].shape[0]
    M = len(levels) - 1
    lev = _i32(np.full(n, M) if level0 is None else level0).copy()
    P, keep = _lf_params(mesh, sym, nsec, lattice, domain, bc_space, bc_time)
    T = substep_table(levels)
    acc = np.zeros((n, 3))
    lib.orc_leapfrog_substeps_f64(ctypes.byref(P), ctypes.addressof(T), M, n,
                                  *[s[k].ctypes.data for k in _STATE_KEYS], lev.ctypes.data,
                                  float(t0), float(dt), int(nsteps), acc.ctypes.data)
    return s, lev, acc


_REF_DIR = os.path.join(_HERE, "_ref")


def ref_pwl_lib():
    """The reference's own pwl_interp_2d_scattered.c built by oracle/ref.mk, or None."""
    p = os.path.join(_REF_DIR, "libref_pwl.so")
    if not os.path.exists(p):
        return None
    L = ctypes.CDLL(p)
    L.pwl_interp_2d_scattered_value.restype = ctypes.c_void_p
    L.pwl_interp_2d_scattered_value.argtypes = [_ci, _vp, _vp, _ci, _vp, _vp, _ci, _vp]
    return L


def ref_efield_lib():
    """The reference's InterpolationScheme built by oracle/ref.mk, or None."""
    p = os.path.join(_REF_DIR, "libref_efield.so")
    if not os.path.exists(p):
        return None
    L = ctypes.CDLL(p)
    L.ref_efield_init.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                  ctypes.c_float, ctypes.c_float]
    L.ref_efield.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]
    return L


# ---------------------------------------------------------------------------
# FMM (oracle/fmm.c)
# ---------------------------------------------------------------------------
lib.orc_fmm_f64.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp, _vp, _ci, _ci, _ci, _vp, _vp]
lib.orc_fmm_taylor.argtypes = [_ci, _vp, _cd, _vp]
lib.orc_fmm_ncoef.argtypes = [_ci]
lib.orc_fmm_root.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp]
lib.orc_fmm_key.restype = ctypes.c_uint64
lib.orc_fmm_key.argtypes = [_cd, _cd, _cd, _vp, _ci]


lib.orc_set_threads.argtypes = [_ci]
lib.orc_get_threads.restype = _ci


def set_threads(n):
    """Threads of the oracle's parallel loops (0: OpenMP's default)."""
    lib.orc_set_threads(int(n))


def get_threads():
    return lib.orc_get_threads()


lib.orc_fmm_set_leaf_stride.argtypes = [_ci]


# FmmType (structures.h:283-288)
FMM_TREECODE, FMM_TREECODE_OUTLIERS, FMM_FMM, FMM_FMM_OUTLIERS = 0, 1, 2, 3
lib.orc_fmm_set_type.argtypes = [_ci]
lib.orc_fmm_outliers.argtypes = [_ci, _vp, _vp, _vp, _vp, _vp]
lib.orc_fmm_outliers.restype = _ci


def fmm_f64(d, p, L, ring=1, leaf_stride=1, fmm_type=FMM_FMM):
    """FP64 FMM accelerations of the engine's algorithm; returns (acc[n,3], stats).
    leaf_stride k > 1 evaluates L2P + P2P on every k-th leaf box only (the
    other rows stay zero), for bounded timing samples.  fmm_type: FMM_FMM
    (M2L/L2L/L2P), FMM_TREECODE (M2P from every level's far boxes), and their
    *_OUTLIERS variants (outliers held out of the tree)."""
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    n = x.shape[0]
    out = np.zeros((n, 3))
    stats = np.zeros(7)
    lib.orc_fmm_set_leaf_stride(int(leaf_stride))
    lib.orc_fmm_set_type(int(fmm_type))
    try:
        lib.orc_fmm_f64(n, x.ctypes.data, y.ctypes.data, z.ctypes.data, q.ctypes.data,
                        st.ctypes.data, sp.ctypes.data, p, L, ring, out.ctypes.data,
                        stats.ctypes.data)
    finally:
        lib.orc_fmm_set_leaf_stride(1)
        lib.orc_fmm_set_type(FMM_FMM)
    return out, {"leaves": int(stats[0]), "p2p_pairs": stats[1], "m2l": stats[2],
                 "far_s": stats[3], "near_s": stats[4], "leaves_evaluated": int(stats[5]),
                 "m2p": stats[6]}


def fmm_outliers(d):
    """Outlier flags (isOutlier as a defined statistic) and their number."""
    x, y, z = _f64(d["x"]), _f64(d["y"]), _f64(d["z"])
    st = _i32(d["status"])
    flags = np.zeros(x.shape[0], np.int32)
    k = lib.orc_fmm_outliers(x.shape[0], x.ctypes.data, y.ctypes.data, z.ctypes.data, st.ctypes.data,
                             flags.ctypes.data)
    return flags, k


def fmm_taylor(p, R, eps2):
    T = np.zeros(lib.orc_fmm_ncoef(p))
    r = _f64(R)
    lib.orc_fmm_taylor(p, r.ctypes.data, float(eps2), T.ctypes.data)
    return T


lib.orc_fmm_counts.argtypes = [_ci, _vp, _vp, _vp, _vp, _ci, _ci, _vp]


def fmm_counts(d, L, ring=2):
    """(M2L box pairs, P2P particle pairs, leaves) of the FMM tree definition."""
    x, y, z = _f64(d["x"]), _f64(d["y"]), _f64(d["z"])
    st = _i32(d["status"])
    out = np.zeros(3)
    lib.orc_fmm_counts(x.shape[0], x.ctypes.data, y.ctypes.data, z.ctypes.data, st.ctypes.data,
                       int(L), int(ring), out.ctypes.data)
    return {"m2l": out[0], "p2p_pairs": out[1], "leaves": int(out[2])}


lib.orc_fmm_set_shell_orders.argtypes = [_vp, _ci]


def fmm_set_shell_orders(orders=None):
    """M2L expansion order per Chebyshev shell of the box offset (index = shell);
    None restores the single order p."""
    if orders is None:
        lib.orc_fmm_set_shell_orders(None, 0)
        return
    a = _i32(orders)
    lib.orc_fmm_set_shell_orders(a.ctypes.data, a.shape[0])


lib.orc_bh_ref2.restype = _ci
lib.orc_bh_ref2.argtypes = [_vp, _vp, _vp, _ci, _ci, _ci, ctypes.c_float, _ci, _ci, _vp, _vp, _ci,
                            _vp, _vp, _vp]
lib.orc_bh_layout.restype = _ci
lib.orc_bh_layout.argtypes = [_vp, _vp, _ci, _ci, _ci, _vp, _vp, _vp, _vp]

BH_INFO_KEYS = ("cells", "root_cells", "max_depth", "buckets", "accepted_cells", "ion_terms",
                "unmatched_root", "nonempty_cells", "opened_cells")


def bh_ref(pos4, vel4, acc4, i0=0, i1=None, theta=0.5, sym="none", lattice=None,
           domain=(0.0, 0.0, 0.0), prism=False, box_dims=None, info=False):
    """Barnes-Hut restatement (the reference's root-cell grid, tree and float
    arithmetic) on AoS float32 arrays (n, 4).  prism: rectangular-prism root
    cells instead of cubic ones (setDomainToggle); box_dims: forced root cell
    counts (setBoxDimVals).  Returns (out (i1-i0, 3) float64, number of cells
    including empty ones), plus a dict of tree and walk counts when info."""
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
    acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
    n = pos4.shape[0]
    i1 = n if i1 is None else i1
    lat = _i32(np.zeros((1, 3)) if lattice is None else lattice)
    dom = _f64(domain)
    out = np.zeros((max(i1 - i0, 0), 3))
    forced = None if box_dims is None else _i32(box_dims)
    inf = np.zeros(len(BH_INFO_KEYS), dtype=np.int64)
    ncell = lib.orc_bh_ref2(pos4.ctypes.data, vel4.ctypes.data, acc4.ctypes.data, n, i0, i1,
                            float(theta), _SYM[sym], lat.shape[0], lat.ctypes.data, dom.ctypes.data,
                            1 if prism else 0, None if forced is None else forced.ctypes.data,
                            out.ctypes.data, inf.ctypes.data)
    if info:
        return out, ncell, dict(zip(BH_INFO_KEYS, (int(v) for v in inf)))
    return out, ncell


def bh_layout(pos4, acc4, i0=0, i1=None, prism=False, box_dims=None):
    """Root-cell grid of the Barnes-Hut tree: (roots (3,) float32, dims (3,)
    float32, cells per axis (3,) int, number of sources)."""
    pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
    acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
    i1 = pos4.shape[0] if i1 is None else i1
    roots = np.zeros(3, dtype=np.float32)
    dims = np.zeros(3, dtype=np.float32)
    nd = np.zeros(3, dtype=np.int32)
    forced = None if box_dims is None else _i32(box_dims)
    ns = lib.orc_bh_layout(pos4.ctypes.data, acc4.ctypes.data, i0, i1, 1 if prism else 0,
                           None if forced is None else forced.ctypes.data, roots.ctypes.data,
                           dims.ctypes.data, nd.ctypes.data)
    return roots, dims, nd, ns


lib.orc_energy_f64.argtypes = [_ci] + [_vp] * 10
lib.orc_field_energy_f64.argtypes = [_cd, _cd, _vp, _ci, _vp, _vp, _vp, _vp, _ci] + [_vp] * 8 + [_ci, _vp]


def energy_f64(d):
    """(kinetic, potential) energy in J of the ions with status 1 or 2 (FP64,
    long double sums)."""
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    vx, vy, vz = _f64(d["vx"]), _f64(d["vy"]), _f64(d["vz"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    out = np.zeros(2)
    lib.orc_energy_f64(x.shape[0], x.ctypes.data, y.ctypes.data, z.ctypes.data, q.ctypes.data,
                       vx.ctypes.data, vy.ctypes.data, vz.ctypes.data, st.ctypes.data,
                       sp.ctypes.data, out.ctypes.data)
    return out[0], out[1]


def field_energy_f64(mesh, p0, p1, q, nsteps=2, fe=None):
    """Field energy increments along the reference's z-then-radial path from
    p0 = (x0, y0, z0) to p1 = (x, y, z)."""
    m = _mesh_arrays(mesh)
    x0, y0, z0 = (_f64(a) for a in p0)
    x, y, z = (_f64(a) for a in p1)
    q = _f64(q)
    n = x.shape[0]
    fe = np.zeros(n) if fe is None else _f64(fe).copy()
    ns_arr = None
    if not np.isscalar(nsteps):
        ns_arr = _i32(nsteps)
    lib.orc_field_energy_f64(float(mesh["r0"]), float(mesh["E0"]), m["xy"].ctypes.data,
                             m["tri"].shape[0], m["tri"].ctypes.data, m["nbr"].ctypes.data,
                             m["er"].ctypes.data, m["ez"].ctypes.data, n, x0.ctypes.data,
                             y0.ctypes.data, z0.ctypes.data, x.ctypes.data, y.ctypes.data,
                             z.ctypes.data, q.ctypes.data,
                             None if ns_arr is None else ns_arr.ctypes.data,
                             0 if ns_arr is not None else int(nsteps), fe.ctypes.data)
    return fe


lib.orc_force_jvp_f64.argtypes = [_ci] + [_vp] * 11
lib.orc_force_vjp_f64.argtypes = [_ci] + [_vp] * 9


def force_jvp_f64(d, dx, dy, dz, dq):
    """Tangent (n, 3) of the FP64 direct acceleration for position/charge tangents."""
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    dx, dy, dz, dq = _f64(dx), _f64(dy), _f64(dz), _f64(dq)
    out = np.zeros((x.shape[0], 3))
    lib.orc_force_jvp_f64(x.shape[0], x.ctypes.data, y.ctypes.data, z.ctypes.data, q.ctypes.data,
                          st.ctypes.data, sp.ctypes.data, dx.ctypes.data, dy.ctypes.data,
                          dz.ctypes.data, dq.ctypes.data, out.ctypes.data)
    return out


def force_vjp_f64(d, b3):
    """Cotangents ((n, 3) positions, (n,) charges) of the FP64 direct
    acceleration for an acceleration cotangent b3 (n, 3)."""
    x, y, z, q = _f64(d["x"]), _f64(d["y"]), _f64(d["z"]), _f64(d["q"])
    st, sp = _i32(d["status"]), _i32(d["species"])
    b3 = _f64(b3)
    xb = np.zeros((x.shape[0], 3))
    qb = np.zeros(x.shape[0])
    lib.orc_force_vjp_f64(x.shape[0], x.ctypes.data, y.ctypes.data, z.ctypes.data, q.ctypes.data,
                          st.ctypes.data, sp.ctypes.data, b3.ctypes.data, xb.ctypes.data,
                          qb.ctypes.data)
    return xb, qb


def rollout_grad_f64(d, dt, nsteps, grad_final):
    """Gradient of an objective of the final state after nsteps DKD steps
    (FP64 direct force, all ions injected, no field or boundaries) with respect
    to the initial positions, velocities and charges, by the discrete adjoint.
    grad_final(x, v) -> (xbar (n, 3), vbar (n, 3)) at the final state.
    Returns (xbar0, vbar0, qbar, final state dict)."""
    cur = {k: np.array(v, copy=True) for k, v in d.items()}
    moving = (cur["status"] == 1) | (cur["status"] == 2)
    kick = cur["status"] == 1
    states = []
    for _ in range(nsteps):
        states.append({k: np.array(v, copy=True) for k, v in cur.items()})
        h = {k: np.array(v, copy=True) for k, v in cur.items()}
        for c, vc in (("x", "vx"), ("y", "vy"), ("z", "vz")):
            h[c] = np.where(moving, cur[c] + 0.5 * dt * cur[vc], cur[c])
        a = direct_f64(h)
        for ax, (c, vc) in enumerate((("x", "vx"), ("y", "vy"), ("z", "vz"))):
            cur[vc] = np.where(kick, cur[vc] + dt * a[:, ax], cur[vc])
            cur[c] = np.where(moving, h[c] + 0.5 * dt * cur[vc], h[c])
    x1 = np.stack([cur["x"], cur["y"], cur["z"]], 1)
    v1 = np.stack([cur["vx"], cur["vy"], cur["vz"]], 1)
    xb, vb = grad_final(x1, v1)
    xb, vb = np.array(xb, float), np.array(vb, float)
    qb = np.zeros(x1.shape[0])
    mv = moving[:, None]
    for s in reversed(states):
        h = {k: np.array(v, copy=True) for k, v in s.items()}
        for c, vc in (("x", "vx"), ("y", "vy"), ("z", "vz")):
            h[c] = np.where(moving, s[c] + 0.5 * dt * s[vc], s[c])
        w = np.where(mv, vb + 0.5 * dt * xb, vb)
        b = np.where(kick[:, None], dt * w, 0.0)
        jx, jq = force_vjp_f64(h, b)
        qb += jq
        xh = xb + jx
        xb = xh
        vb = np.where(mv, w + 0.5 * dt * xh, w)
    return xb, vb, qb, cur


# ---- the reference's rotation-based spherical-harmonic FMM (oracle/fmm_sph.c) ----

lib.orc_domain_bounds.restype = _ci
lib.orc_domain_bounds.argtypes = [_vp, _vp, _ci, _ci, _vp, _ci, _vp, _vp, _vp]
lib.orc_fmm_sph.restype = _ci
lib.orc_fmm_sph.argtypes = [_vp, _vp, _vp, _ci, _ci, _ci, _ci, _ci, _ci, _vp, _vp]
lib.orc_sph_p2m.argtypes = [_ci, _ci, _vp, ctypes.c_double, _vp]
lib.orc_sph_m2m.argtypes = [_ci, _ci, _vp, _vp]
lib.orc_sph_m2l.argtypes = [_ci, _ci, _ci, _ci, ctypes.c_double, _vp, _vp]
lib.orc_sph_l2l.argtypes = [_ci, _ci, _vp, _vp]
lib.orc_sph_l2p.argtypes = [_ci, _vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, _vp]
lib.orc_sph_rotate.restype = _ci
lib.orc_sph_rotate.argtypes = [_ci, _ci, _ci, _ci, _ci, _vp, _vp]
lib.orc_sph_set_cr_trig.argtypes = [_ci]
lib.orc_sph_table.restype = ctypes.c_double
lib.orc_sph_table.argtypes = [_ci, _ci]


def domain_bounds(pos4, acc4, i0=0, i1=None, rotational=False):
    """computeDomainBounds as the FMM calls it (computeDomainBoundsRotations with
    rotational=True): (boxMin[3], dims[3], boxDimVals[3], nsrc)."""
    pos4, acc4 = _f32(pos4), _f32(acc4)
    i1 = pos4.shape[0] if i1 is None else i1
    bm, dm, nd = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.int32)
    ns = lib.orc_domain_bounds(pos4.ctypes.data, acc4.ctypes.data, i0, i1, None, 1 if rotational else 0,
                               bm.ctypes.data, dm.ctypes.data, nd.ctypes.data)
    return bm, dm, nd, ns


def fmm_sph(pos4, vel4, acc4, p=10, max_level=-2, i0=0, i1=None, cr_trig=False, rotational=False):
    """The reference FMM's operators (rotation-based spherical harmonics) over the
    reference's domain grid and leaf placement, in FP64; returns (acc[i1-i0,3], info).
    max_level: setMaxLevel's value (-2: setOptimumLevel).  rotational: the
    reference's rotational FMM (four quadrant images about z).  Raises on the
    cases the restatement does not cover (sources outside the grid, maxLevel < 2,
    a rotational grid with more than one domain in x or y).
    cr_trig: correctly rounded float acos / atan in cart2sph_exact instead of
    this host's acosf / atanf (which the reference calls)."""
    pos4, vel4, acc4 = _f32(pos4), _f32(vel4), _f32(acc4)
    n = pos4.shape[0]
    i1 = n if i1 is None else i1
    out = np.zeros((i1 - i0, 3))
    info = np.zeros(12, np.int64)
    lib.orc_sph_set_cr_trig(1 if cr_trig else 0)
    try:
        rc = lib.orc_fmm_sph(pos4.ctypes.data, vel4.ctypes.data, acc4.ctypes.data, n, i0, i1, p,
                             max_level, 1 if rotational else 0, out.ctypes.data, info.ctypes.data)
    finally:
        lib.orc_sph_set_cr_trig(0)
    if rc != 0:
        raise ValueError(f"orc_fmm_sph: unsupported case ({rc})")
    keys = ["max_level", "k", "nd0", "nd1", "nd2", "leaves", "m2l_pairs", "p2p_pairs", "sources",
            "boxes", "p2p_ion_pairs"]
    return out, {k: int(info[i]) for i, k in enumerate(keys)}


lib.orc_fmm_sph_trans.restype = _ci
lib.orc_fmm_sph_trans.argtypes = [_vp, _vp, _vp, _ci, _ci, _ci, _ci, _ci, ctypes.c_float, _ci, _vp, _vp]


def fmm_sph_trans(pos4, vel4, acc4, cell, p=10, max_level=-2, i0=0, i1=None, near=True, cr_trig=False):
    """The reference's translational FMM (ternary tree over one periodic cubic
    cell of side `cell` on all three axes, lists wrapped periodically) with the
    operators of fmm_sph; returns (acc[i1-i0,3], info) like fmm_sph.  Every
    target sees the sources of the 27 cells at shifts {-1, 0, 1}^3.  near=False
    gives the reference as it runs: its translational P2P passes the thread's
    box range as the ion range and adds nothing (fmm_cpukernel_translational.
    cpp:326-332).  Raises for a source outside the cell, maxLevel < 1 or > 6."""
    pos4, vel4, acc4 = _f32(pos4), _f32(vel4), _f32(acc4)
    n = pos4.shape[0]
    i1 = n if i1 is None else i1
    out = np.zeros((i1 - i0, 3))
    info = np.zeros(12, np.int64)
    lib.orc_sph_set_cr_trig(1 if cr_trig else 0)
    try:
        rc = lib.orc_fmm_sph_trans(pos4.ctypes.data, vel4.ctypes.data, acc4.ctypes.data, n, i0, i1, p,
                                   max_level, float(cell), 1 if near else 0, out.ctypes.data, info.ctypes.data)
    finally:
        lib.orc_sph_set_cr_trig(0)
    if rc != 0:
        raise ValueError(f"orc_fmm_sph_trans: unsupported case ({rc})")
    keys = ["max_level", "k", "nd0", "nd1", "nd2", "leaves", "m2l_pairs", "p2p_pairs", "sources",
            "boxes", "p2p_ion_pairs"]
    return out, {k: int(info[i]) for i, k in enumerate(keys)}


def _sph_nc(p):
    return p * (p + 1) // 2


def sph_p2m(p, xyzq, h):
    """Multipole (complex, n(n+1)/2 + m order) of particles at xyzq[:, :3] from the centre."""
    a = _f64(np.asarray(xyzq).reshape(-1, 4))
    M = np.zeros(2 * _sph_nc(p))
    lib.orc_sph_p2m(p, a.shape[0], a.ctypes.data, float(h), M.ctypes.data)
    return M[0::2] + 1j * M[1::2]


def _cz(v):
    v = np.asarray(v, np.complex128)
    out = np.zeros(2 * v.shape[0])
    out[0::2], out[1::2] = v.real, v.imag
    return out


def sph_m2m(p, octant, Mc, Mp=None):
    a = _cz(Mc)
    b = _cz(np.zeros(_sph_nc(p)) if Mp is None else Mp)
    lib.orc_sph_m2m(p, octant, a.ctypes.data, b.ctypes.data)
    return b[0::2] + 1j * b[1::2]


def sph_m2l(p, d, h, M, L=None):
    a = _cz(M)
    b = _cz(np.zeros(_sph_nc(p)) if L is None else L)
    lib.orc_sph_m2l(p, int(d[0]), int(d[1]), int(d[2]), float(h), a.ctypes.data, b.ctypes.data)
    return b[0::2] + 1j * b[1::2]


def sph_l2l(p, octant, Lp):
    a = _cz(Lp)
    b = _cz(np.zeros(_sph_nc(p)))
    lib.orc_sph_l2l(p, octant, a.ctypes.data, b.ctypes.data)
    return b[0::2] + 1j * b[1::2]


def sph_l2p(p, L, dxyz, h):
    """The field vector the reference's L2P forms before scaling by -K q / m."""
    a = _cz(L)
    f = np.zeros(3)
    lib.orc_sph_l2p(p, a.ctypes.data, float(dxyz[0]), float(dxyz[1]), float(dxyz[2]), float(h),
                    f.ctypes.data)
    return f


def sph_rotate(p, d, C, backward=False):
    a = _cz(C)
    b = np.zeros_like(a)
    if lib.orc_sph_rotate(p, int(d[0]), int(d[1]), int(d[2]), 1 if backward else 0, a.ctypes.data,
                          b.ctypes.data) != 0:
        raise ValueError("direction must have gcd 1")
    return b[0::2] + 1j * b[1::2]


def sph_table(which, index):
    return lib.orc_sph_table({"factorial": 0, "anm": 1, "Anm": 2, "Ynm": 3}[which], int(index))
