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
