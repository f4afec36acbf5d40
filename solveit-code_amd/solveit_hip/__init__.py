"""ctypes binding of libsolveit_hip.so (the MI355X force engine's C ABI).

This is the harness-side view of include/solveit_hip.h.  Importing it loads the
in-tree library and fails loudly when it has not been built: there is no CPU
fallback anywhere in the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsolveit_hip.so")

SHP_PREC_REF = 0
SHP_PREC_FP64 = 1
SHP_PREC_FP32C = 2

SHP_SYM_NONE = 0
SHP_SYM_ROTATIONAL = 1
SHP_SYM_TRANSLATIONAL = 2

# ParticleStatus (structures.h:206-211) and ParticleSpecies (:224-233)
STATUS_INBOUND = 1
STATUS_AFFECTS = 2
STATUS_OUT = 0
STATUS_CRASHED = -1
MONOMER_POS, DIMER_POS, TRIMER_POS, DROPLET, NEUTRAL, MONOMER_NEG, DIMER_NEG, TRIMER_NEG = range(8)


class ShpError(RuntimeError):
    pass


class Symmetry(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int),
        ("nsec", ctypes.c_int),
        ("nimg", ctypes.c_int),
        ("lattice", ctypes.POINTER(ctypes.c_int)),
        ("domain", ctypes.c_double * 3),
    ]


class FieldMesh(ctypes.Structure):
    _fields_ = [
        ("nnodes", ctypes.c_int),
        ("xy", ctypes.c_void_p),
        ("er", ctypes.c_void_p),
        ("ez", ctypes.c_void_p),
        ("ntri", ctypes.c_int),
        ("tri", ctypes.c_void_p),
        ("nbr", ctypes.c_void_p),
        ("r0", ctypes.c_double),
        ("E0", ctypes.c_double),
    ]


class BhOptions(ctypes.Structure):
    _fields_ = [
        ("domain_toggle", ctypes.c_int),
        ("force_box_dims", ctypes.c_int),
        ("box_dims", ctypes.c_int * 3),
        ("subcell_block", ctypes.c_int),
        ("outlier_toggle", ctypes.c_int),
        ("walk", ctypes.c_int),
    ]


BH_STAT_KEYS = ("root_cells", "nodes", "max_depth", "buckets", "accepted_cells", "ion_terms",
                "unmatched_root", "nx", "ny", "nz", "root_x", "root_y", "root_z", "dim_x", "dim_y",
                "dim_z", "opened_cells")


class SubstepLevel(ctypes.Structure):
    _fields_ = [("bc_type", ctypes.c_int), ("bc_space", ctypes.c_double * 4), ("bc_time", ctypes.c_double)]


def make_substep_levels(levels):
    """levels: list of (bc_type, (bx, by, bz, brho), bt), one per level 0..M."""
    arr = (SubstepLevel * len(levels))()
    for k, (typ, sp, tm) in enumerate(levels):
        arr[k].bc_type = int(typ)
        for c in range(4):
            arr[k].bc_space[c] = float(sp[c])
        arr[k].bc_time = float(tm)
    return arr


class State(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("z", ctypes.c_void_p),
        ("vx", ctypes.c_void_p),
        ("vy", ctypes.c_void_p),
        ("vz", ctypes.c_void_p),
        ("q", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("species", ctypes.c_void_p),
        ("t_inj", ctypes.c_void_p),
    ]


class StepOpts(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("sym", Symmetry),
        ("use_efield", ctypes.c_int),
        ("bc_space", ctypes.c_double * 4),
        ("bc_time", ctypes.c_double),
        ("fmm_order", ctypes.c_int),
        ("leaf_size", ctypes.c_int),
        ("theta", ctypes.c_double),
        ("fmm_ring", ctypes.c_int),
        ("leaf_level", ctypes.c_int),
    ]


SHP_METHOD_DIRECT = 0
SHP_METHOD_FMM = 1
SHP_METHOD_BH = 2


class Particles(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int),
        ("x", ctypes.c_void_p),
        ("y", ctypes.c_void_p),
        ("z", ctypes.c_void_p),
        ("q", ctypes.c_void_p),
        ("status", ctypes.c_void_p),
        ("species", ctypes.c_void_p),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() "
            "(hipcc --offload-arch=gfx950) first")
    # torch ships its own HIP/HSA runtime (soname libamdhip64.so.7 under a
    # different file name).  Loading torch first makes the dynamic loader bind
    # this library's libamdhip64.so.7 / libhsa-runtime64.so.1 to those already
    # loaded copies, so one process never initialises two HSA runtimes.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    c_int, c_double, c_void_p = ctypes.c_int, ctypes.c_double, ctypes.c_void_p
    fp = ctypes.POINTER(ctypes.c_float)
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "shp_version": (c_int, []),
        "shp_last_error": (ctypes.c_char_p, []),
        "shp_create": (c_int, [c_int, ctypes.POINTER(c_void_p)]),
        "shp_destroy": (c_int, [c_void_p]),
        "shp_sync": (c_int, [c_void_p]),
        "shp_stream": (c_void_p, [c_void_p]),
        "shp_k_const": (c_double, []),
        "shp_softening": (c_double, []),
        "shp_species_mass": (c_double, [c_int]),
        "shp_direct_host": (c_int, [c_void_p, fp, fp, fp, c_int, c_int, c_int, c_int, c_int,
                                    c_int, ctypes.POINTER(Symmetry), dp]),
        "shp_direct_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), c_int,
                                   ctypes.POINTER(Symmetry), c_void_p, c_void_p, c_void_p]),
        "shp_last_kernel_ms": (c_double, [c_void_p]),
        "shp_efield_setup": (c_int, [c_void_p, ctypes.POINTER(FieldMesh)]),
        "shp_efield_dev": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p]),
        "shp_leapfrog_run": (c_int, [c_void_p, ctypes.POINTER(State), ctypes.POINTER(StepOpts),
                                     c_double, c_double, c_int, c_void_p, c_void_p, c_void_p]),
        "shp_last_update_ms": (c_double, [c_void_p]),
        "shp_fmm_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), c_int, c_int, c_int, c_int,
                                c_void_p, c_void_p, c_void_p]),
        "shp_fmm_last_leaf_level": (c_int, [c_void_p]),
        "shp_fmm_sym_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), ctypes.POINTER(Symmetry),
                                    c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
        "shp_fmm_stats": (c_int, [c_void_p, dp]),
        "shp_fmm_timing": (c_int, [c_void_p, dp, c_int]),
        "shp_fmm_level_estimates": (c_int, [c_void_p, dp, c_int]),
        "shp_fmm_shell_orders": (c_int, [c_int, c_int, c_int, ctypes.POINTER(c_int)]),
        "shp_compare_accel": (c_int, [c_void_p, c_int] + [c_void_p] * 7 + [dp]),
        "shp_energy_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), c_void_p, c_void_p,
                                   c_void_p, dp]),
        "shp_field_energy_dev": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
        "shp_force_jvp_dev": (c_int, [c_void_p, ctypes.POINTER(Particles)] + [c_void_p] * 7),
        "shp_force_vjp_dev": (c_int, [c_void_p, ctypes.POINTER(Particles)] + [c_void_p] * 7
                              + [c_int]),
        "shp_leapfrog_step_vjp_dev": (c_int, [c_void_p, ctypes.POINTER(State), c_double]
                                      + [c_void_p] * 13),
        "shp_batch_create": (c_int, [c_void_p, c_int, ctypes.POINTER(c_int), c_int,
                                     ctypes.POINTER(c_void_p)]),
        "shp_batch_destroy": (c_int, [c_void_p]),
        "shp_batch_tune": (c_int, [c_void_p, ctypes.POINTER(Particles), dp,
                                   ctypes.POINTER(c_int)]),
        "shp_batch_set_config": (c_int, [c_void_p, c_int]),
        "shp_batch_forces": (c_int, [c_void_p, ctypes.POINTER(Particles), c_void_p, c_void_p,
                                     c_void_p]),
        "shp_batch_leapfrog": (c_int, [c_void_p, ctypes.POINTER(State), ctypes.POINTER(StepOpts),
                                       c_double, c_double, c_int, c_int, c_void_p, c_void_p,
                                       c_void_p]),
        "shp_batch_step_vjp": (c_int, [c_void_p, ctypes.POINTER(State), c_double, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p]),
        "shp_bh_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), ctypes.c_float,
                               ctypes.POINTER(Symmetry), c_void_p, c_void_p, c_void_p]),
        "shp_bh_host": (c_int, [c_void_p, fp, fp, fp, c_int, c_int, c_int, ctypes.c_float,
                                ctypes.POINTER(Symmetry), dp]),
        "shp_bh_set_options": (c_int, [c_void_p, ctypes.POINTER(BhOptions)]),
        "shp_field_mesh_load_csv": (c_int, [ctypes.c_char_p, ctypes.c_char_p, c_double, c_double,
                                            ctypes.POINTER(FieldMesh)]),
        "shp_field_mesh_free": (None, [ctypes.POINTER(FieldMesh)]),
        "shp_state_upload_host": (c_int, [c_void_p, fp, fp, fp, fp, c_int]),
        "shp_leapfrog_substeps_run": (c_int, [c_void_p, ctypes.POINTER(State), ctypes.POINTER(StepOpts),
                                              ctypes.POINTER(StepOpts), c_void_p, c_int, c_void_p, c_double,
                                              c_double, c_int, c_void_p, c_void_p, c_void_p]),
        "shp_state_run": (c_int, [c_void_p, ctypes.POINTER(StepOpts), c_double, c_double, c_int]),
        "shp_state_run_substeps": (c_int, [c_void_p, ctypes.POINTER(StepOpts), ctypes.POINTER(StepOpts), c_void_p,
                                           c_int, c_int, c_double, c_double, c_int]),
        "shp_state_download_host": (c_int, [c_void_p, fp, fp, fp]),
        "shp_fmm_set_type": (c_int, [c_void_p, c_int]),
        "shp_fmm_outliers_dev": (c_int, [c_void_p, ctypes.POINTER(Particles), c_void_p, ctypes.POINTER(c_int)]),
        "shp_state_save": (c_int, [c_void_p, c_int, c_int, c_int, ctypes.c_float, c_void_p]),
        "shp_state_save_full": (c_int, [c_void_p, c_int, c_int, ctypes.c_float, c_void_p]),
        "shp_state_flush": (c_int, [c_void_p]),
        "shp_state_events": (c_int, [c_void_p, c_void_p, ctypes.POINTER(c_int)]),
        "shp_state_fields_host": (c_int, [c_void_p, fp, fp]),
        "shp_fmm_host": (c_int, [c_void_p, fp, fp, fp, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                 ctypes.POINTER(Symmetry), dp]),
        "shp_direct_dev_noalloc": (c_int, [c_void_p, ctypes.POINTER(Particles), c_int,
                                           ctypes.POINTER(Symmetry), c_void_p, c_void_p, c_void_p]),
        "shp_bh_stats": (c_int, [c_void_p, dp]),
        "shp_fmm_sph_host": (c_int, [c_void_p, fp, fp, fp, c_int, c_int, c_int, c_int, c_int,
                                     ctypes.POINTER(Symmetry), dp]),
        "shp_fmm_sph_stats": (c_int, [c_void_p, dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib, sig


lib, _SIGNATURES = _load()
EXPORTED = tuple(_SIGNATURES)


def _check(rc, what):
    if rc != 0:
        raise ShpError(f"{what} failed ({rc}): {lib.shp_last_error().decode()}")


def k_const():
    return lib.shp_k_const()


def softening():
    return lib.shp_softening()


def _torch_ready(device):
    """Device calls run on the engine's own stream: wait for torch's current
    stream first so its writes to the inputs and outputs have landed.  Every
    wrapper below synchronizes the engine's stream before returning tensors."""
    import torch
    torch.cuda.current_stream(device).synchronize()


def fmm_shell_orders(order, ring=2, rotational=False):
    """M2L orders on the far-band shells 3, 4, 5 for an FMM of this order
    (0, 0, 0: the table-driven M2L at the single order)."""
    out = (ctypes.c_int * 3)()
    _check(lib.shp_fmm_shell_orders(int(order), int(ring),
                                    SHP_SYM_ROTATIONAL if rotational else SHP_SYM_NONE, out),
           "shp_fmm_shell_orders")
    return tuple(out)


SHP_FMM_TREECODE = 0
SHP_FMM_TREECODE_OUTLIERS = 1
SHP_FMM_FMM = 2
SHP_FMM_FMM_OUTLIERS = 3

SHP_SAVE_ORIGINAL = 0
SHP_SAVE_NEW_ORIGINAL = 1

_libc = ctypes.CDLL(None)
_libc.fopen.restype = ctypes.c_void_p
_libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
_libc.fclose.argtypes = [ctypes.c_void_p]
_libc.fclose.restype = ctypes.c_int


class CFile:
    """A C stdio FILE* for the writer entry points (fopen / fclose of the C
    library the engine links)."""

    def __init__(self, path, mode="wb"):
        self.ptr = _libc.fopen(str(path).encode(), mode.encode())
        if not self.ptr:
            raise OSError(f"fopen({path!r}, {mode!r}) failed")

    def close(self):
        if self.ptr:
            _libc.fclose(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def load_field_mesh_csv(fields_path, triangles_path, r0, E0):
    """The reference's Fields_*.csv / Triangles_*.csv read by the library's loader;
    returns the mesh dict (xy, er, ez, tri, nbr, r0, E0) as numpy copies."""
    m = FieldMesh()
    _check(lib.shp_field_mesh_load_csv(str(fields_path).encode(), str(triangles_path).encode(), float(r0),
                                       float(E0), ctypes.byref(m)), "shp_field_mesh_load_csv")
    try:
        nd, nt = m.nnodes, m.ntri

        def arr(ptr, ctype, count):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), shape=(count,)).copy()

        out = {"xy": arr(m.xy, ctypes.c_double, 2 * nd).reshape(nd, 2),
               "er": arr(m.er, ctypes.c_double, nd), "ez": arr(m.ez, ctypes.c_double, nd),
               "tri": arr(m.tri, ctypes.c_int, 3 * nt).reshape(nt, 3),
               "nbr": arr(m.nbr, ctypes.c_int, 3 * nt).reshape(nt, 3), "r0": m.r0, "E0": m.E0}
    finally:
        lib.shp_field_mesh_free(ctypes.byref(m))
    return out


def make_symmetry(kind="none", nsec=4, lattice=None, domain=(0.0, 0.0, 0.0)):
    """Build a Symmetry struct; keeps the lattice array alive on the struct."""
    s = Symmetry()
    s.type = {"none": SHP_SYM_NONE, "rotational": SHP_SYM_ROTATIONAL,
              "translational": SHP_SYM_TRANSLATIONAL}[kind]
    s.nsec = int(nsec)
    s.nimg = 0
    s.lattice = None
    if kind == "translational":
        lat = np.ascontiguousarray(np.asarray(lattice, dtype=np.int32).reshape(-1, 3))
        s._lat = lat
        s.nimg = lat.shape[0]
        s.lattice = lat.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    for c in range(3):
        s.domain[c] = float(domain[c])
    return s


class Engine:
    """One HIP context (device, stream, workspace) of the force engine."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib.shp_create(int(device), ctypes.byref(h)), "shp_create")
        self.h = h
        self.device = device

    def close(self):
        if self.h:
            lib.shp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        return lib.shp_stream(self.h)

    def sync(self):
        _check(lib.shp_sync(self.h), "shp_sync")

    def last_kernel_ms(self):
        return lib.shp_last_kernel_ms(self.h)

    def direct_host(self, pos4, vel4, acc4, i0=None, i1=None, j0=None, j1=None,
                    precision=SHP_PREC_REF, symmetry=None):
        """Reference-contract direct sum on host AoS float32 arrays of shape (n, 4).

        Returns the (i1-i0, 3) float64 acceleration before the reference's float
        rounding (zero rows for non-targets)."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
        acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
        n = pos4.shape[0]
        i0 = 0 if i0 is None else i0
        i1 = n if i1 is None else i1
        j0 = 0 if j0 is None else j0
        j1 = n if j1 is None else j1
        out = np.zeros((max(i1 - i0, 0), 3), dtype=np.float64)
        fp = ctypes.POINTER(ctypes.c_float)
        sym = symmetry if symmetry is not None else make_symmetry("none")
        rc = lib.shp_direct_host(self.h, pos4.ctypes.data_as(fp), vel4.ctypes.data_as(fp),
                                 acc4.ctypes.data_as(fp), n, i0, i1, j0, j1, precision,
                                 ctypes.byref(sym),
                                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        _check(rc, "shp_direct_host")
        return out

    def direct_dev(self, state, precision=SHP_PREC_FP64, symmetry=None, out=None):
        """Direct sum on a DeviceState (torch tensors on the GPU).  Returns
        (ax, ay, az) float64 device tensors; stream-ordered on the engine's stream."""
        import torch
        n = state.n
        if out is None:
            out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device)
                        for _ in range(3))
        sym = symmetry if symmetry is not None else make_symmetry("none")
        p = state.as_struct()
        _torch_ready(state.x.device)
        rc = lib.shp_direct_dev(self.h, ctypes.byref(p), precision, ctypes.byref(sym),
                                out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
        _check(rc, "shp_direct_dev")
        self.sync()
        return out

    def bh_host(self, pos4, vel4, acc4, i0=None, i1=None, theta=0.5, symmetry=None):
        """Barnes-Hut (reference FP32 arithmetic) on host AoS float32 arrays of
        shape (n, 4); ions [i0, i1) are sources and targets.  Returns the
        (i1-i0, 3) float64 copy of the float acceleration (zero rows for
        non-targets)."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
        acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
        n = pos4.shape[0]
        i0 = 0 if i0 is None else i0
        i1 = n if i1 is None else i1
        out = np.zeros((max(i1 - i0, 0), 3), dtype=np.float64)
        fp = ctypes.POINTER(ctypes.c_float)
        sym = symmetry if symmetry is not None else make_symmetry("none")
        rc = lib.shp_bh_host(self.h, pos4.ctypes.data_as(fp), vel4.ctypes.data_as(fp),
                             acc4.ctypes.data_as(fp), n, i0, i1, float(theta), ctypes.byref(sym),
                             out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        _check(rc, "shp_bh_host")
        return out

    def fmm_host(self, pos4, vel4, acc4, i0=None, i1=None, order=10, leaf_level=0, leaf_size=20,
                 ring=2, symmetry=None):
        """FMM on host AoS float32 arrays of shape (n, 4); ions [i0, i1) are sources
        and targets.  Returns the (i1-i0, 3) float64 acceleration (zero rows for
        non-targets)."""
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
        acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
        n = pos4.shape[0]
        i0 = 0 if i0 is None else i0
        i1 = n if i1 is None else i1
        out = np.zeros((max(i1 - i0, 0), 3), dtype=np.float64)
        fp = ctypes.POINTER(ctypes.c_float)
        sym = symmetry if symmetry is not None else make_symmetry("none")
        rc = lib.shp_fmm_host(self.h, pos4.ctypes.data_as(fp), vel4.ctypes.data_as(fp),
                              acc4.ctypes.data_as(fp), n, i0, i1, int(order), int(leaf_level),
                              int(leaf_size), int(ring), ctypes.byref(sym),
                              out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        _check(rc, "shp_fmm_host")
        return out

    def fmm_sph_host(self, pos4, vel4, acc4, i0=None, i1=None, order=10, max_level=-2, symmetry=None):
        """The reference FMM's own operators (rotation-based spherical harmonics,
        FP64) on host AoS float32 arrays; returns the (i1-i0, 3) float64
        acceleration.  max_level -2: the reference's setOptimumLevel choice.
        symmetry: None, make_symmetry("rotational", 4) (the reference's
        rotational FMM) or make_symmetry("translational", lattice=the 27 points
        {-1, 0, 1}^3, domain=(L, L, L)) (its translational FMM over the periodic
        cell [-L/2, L/2)^3, P2P included)."""
        sym = symmetry if symmetry is not None else make_symmetry("none")
        pos4 = np.ascontiguousarray(pos4, dtype=np.float32)
        vel4 = np.ascontiguousarray(vel4, dtype=np.float32)
        acc4 = np.ascontiguousarray(acc4, dtype=np.float32)
        n = pos4.shape[0]
        i0 = 0 if i0 is None else i0
        i1 = n if i1 is None else i1
        out = np.zeros((max(i1 - i0, 0), 3), dtype=np.float64)
        fp = ctypes.POINTER(ctypes.c_float)
        rc = lib.shp_fmm_sph_host(self.h, pos4.ctypes.data_as(fp), vel4.ctypes.data_as(fp),
                                  acc4.ctypes.data_as(fp), n, i0, i1, int(order), int(max_level),
                                  ctypes.byref(sym), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        _check(rc, "shp_fmm_sph_host")
        return out

    def fmm_sph_stats(self):
        """Summary of the last fmm_sph_host call."""
        out = np.zeros(12)
        _check(lib.shp_fmm_sph_stats(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "shp_fmm_sph_stats")
        keys = ["max_level", "k", "nd0", "nd1", "nd2", "leaves", "m2l_pairs", "p2p_pairs", "sources",
                "boxes", "p2p_ion_pairs"]
        return {k: int(out[i]) for i, k in enumerate(keys)}

    def state_upload_host(self, pos4, vel4, acc4, t_inj):
        """Copies the reference's host arrays into the engine's resident FP64 state."""
        fp = ctypes.POINTER(ctypes.c_float)
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (pos4, vel4, acc4, t_inj)]
        _check(lib.shp_state_upload_host(self.h, *(a.ctypes.data_as(fp) for a in arrs), arrs[0].shape[0]),
               "shp_state_upload_host")

    def state_run(self, opts, t0, dt, nsteps):
        _check(lib.shp_state_run(self.h, ctypes.byref(opts), float(t0), float(dt), int(nsteps)),
               "shp_state_run")

    def state_run_substeps(self, opts, levels, t0, dt, nsteps, reset=False, sub=None):
        table = make_substep_levels(levels)
        _check(lib.shp_state_run_substeps(self.h, ctypes.byref(opts), ctypes.byref(sub) if sub is not None else None,
                                          ctypes.addressof(table), len(levels) - 1, 1 if reset else 0, float(t0),
                                          float(dt), int(nsteps)), "shp_state_run_substeps")

    def state_download_host(self, pos4, vel4, acc4):
        """Writes the resident state back into float32 (n, 4) arrays (in place)."""
        fp = ctypes.POINTER(ctypes.c_float)
        for a in (pos4, vel4, acc4):
            assert a.dtype == np.float32 and a.flags.c_contiguous
        _check(lib.shp_state_download_host(self.h, pos4.ctypes.data_as(fp), vel4.ctypes.data_as(fp),
                                           acc4.ctypes.data_as(fp)), "shp_state_download_host")

    def state_save(self, cfile, fmt, first, last, time):
        """Integrator::saveData of the resident state into an open CFile (asynchronous)."""
        _check(lib.shp_state_save(self.h, int(fmt), int(first), int(last), float(time), cfile.ptr),
               "shp_state_save")

    def state_save_full(self, cfile, fmt, last, time):
        """Integrator::saveFullCurrentConditions of the resident state (asynchronous)."""
        _check(lib.shp_state_save_full(self.h, int(fmt), int(last), float(time), cfile.ptr),
               "shp_state_save_full")

    def state_flush(self):
        _check(lib.shp_state_flush(self.h), "shp_state_flush")

    def state_events(self, cfile):
        """Writes the queued boundary-condition events; returns their number."""
        ne = ctypes.c_int(0)
        _check(lib.shp_state_events(self.h, cfile.ptr, ctypes.byref(ne)), "shp_state_events")
        return ne.value

    def state_fields_host(self, n):
        """Efield1 and Efield2 of the resident ions as float32 (n, 3) arrays."""
        fp = ctypes.POINTER(ctypes.c_float)
        e1 = np.zeros((n, 3), np.float32)
        e2 = np.zeros((n, 3), np.float32)
        _check(lib.shp_state_fields_host(self.h, e1.ctypes.data_as(fp), e2.ctypes.data_as(fp)),
               "shp_state_fields_host")
        return e1, e2

    def bh_set_options(self, prism=False, box_dims=None, subcell_block=2, outlier_toggle=0, walk=0):
        """Root-layout options of later Barnes-Hut calls (BarnesHutForceCalculation's
        setDomainToggle / setBoxDimVals / setSubcellBlockSize / setOutlierToggle), and the
        GPU walk (0 by size, 1 per lane, 2 wave-coherent)."""
        o = BhOptions()
        o.domain_toggle = 1 if prism else 0
        o.force_box_dims = 0 if box_dims is None else 1
        for c in range(3):
            o.box_dims[c] = 1 if box_dims is None else int(box_dims[c])
        o.subcell_block = int(subcell_block)
        o.outlier_toggle = int(outlier_toggle)
        o.walk = int(walk)
        _check(lib.shp_bh_set_options(self.h, ctypes.byref(o)), "shp_bh_set_options")

    def bh_stats(self):
        """Tree and walk counts of the last Barnes-Hut evaluation (dict)."""
        out = (ctypes.c_double * len(BH_STAT_KEYS))()
        _check(lib.shp_bh_stats(self.h, out), "shp_bh_stats")
        return dict(zip(BH_STAT_KEYS, list(out)))

    def bh_dev(self, state, theta=0.5, symmetry=None, out=None):
        """Barnes-Hut on a DeviceState; returns (ax, ay, az) float64 device tensors."""
        import torch
        n = state.n
        if out is None:
            out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device)
                        for _ in range(3))
        sym = symmetry if symmetry is not None else make_symmetry("none")
        p = state.as_struct()
        _torch_ready(state.x.device)
        rc = lib.shp_bh_dev(self.h, ctypes.byref(p), float(theta), ctypes.byref(sym),
                            out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
        _check(rc, "shp_bh_dev")
        self.sync()
        return out

    def energy(self, state):
        """(kinetic, potential) energy in J of a DeviceState's ions with status 1 or 2."""
        out = (ctypes.c_double * 2)()
        p = state.as_struct()
        _torch_ready(state.x.device)
        rc = lib.shp_energy_dev(self.h, ctypes.byref(p), state.vx.data_ptr(),
                                state.vy.data_ptr(), state.vz.data_ptr(), out)
        _check(rc, "shp_energy_dev")
        return out[0], out[1]

    def field_energy(self, x0, y0, z0, x, y, z, q, nsteps=2, fe=None):
        """Adds the field energy along the reference's z-then-radial path to `fe`
        (float64 device tensors; nsteps an int or an int32 tensor); returns fe."""
        import torch
        if fe is None:
            fe = torch.zeros_like(x)
        ns_arr = nsteps.data_ptr() if hasattr(nsteps, "data_ptr") else None
        ns = 0 if ns_arr is not None else int(nsteps)
        _torch_ready(x.device)
        rc = lib.shp_field_energy_dev(self.h, x.shape[0], x0.data_ptr(), y0.data_ptr(),
                                      z0.data_ptr(), x.data_ptr(), y.data_ptr(), z.data_ptr(),
                                      q.data_ptr(), ns_arr, ns, fe.data_ptr())
        _check(rc, "shp_field_energy_dev")
        self.sync()
        return fe

    def force_jvp(self, state, dx=None, dy=None, dz=None, dq=None):
        """Tangent (ax, ay, az) of the FP64 direct acceleration (device tensors)."""
        import torch
        n = state.n
        out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device) for _ in range(3))
        p = state.as_struct()
        ptr = lambda t: None if t is None else t.data_ptr()
        _torch_ready(state.x.device)
        rc = lib.shp_force_jvp_dev(self.h, ctypes.byref(p), ptr(dx), ptr(dy), ptr(dz), ptr(dq),
                                   out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr())
        _check(rc, "shp_force_jvp_dev")
        self.sync()
        return out

    def force_vjp(self, state, bx, by, bz):
        """Cotangents ((xb, yb, zb), qb) for an acceleration cotangent."""
        import torch
        n = state.n
        out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device) for _ in range(4))
        p = state.as_struct()
        _torch_ready(state.x.device)
        rc = lib.shp_force_vjp_dev(self.h, ctypes.byref(p), bx.data_ptr(), by.data_ptr(),
                                   bz.data_ptr(), out[0].data_ptr(), out[1].data_ptr(),
                                   out[2].data_ptr(), out[3].data_ptr(), 0)
        _check(rc, "shp_force_vjp_dev")
        self.sync()
        return out[:3], out[3]

    def leapfrog_step_vjp(self, state, dt, xb1, vb1, qb):
        """VJP of one DKD step from `state` (start of the step): xb1, vb1 are
        3-tuples of device tensors; returns (xb0, vb0) and adds into qb."""
        import torch
        n = state.n
        out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device) for _ in range(6))
        s = state.as_state()
        _torch_ready(state.x.device)
        rc = lib.shp_leapfrog_step_vjp_dev(self.h, ctypes.byref(s), float(dt),
                                           *(t.data_ptr() for t in xb1),
                                           *(t.data_ptr() for t in vb1),
                                           *(t.data_ptr() for t in out), qb.data_ptr())
        _check(rc, "shp_leapfrog_step_vjp_dev")
        self.sync()
        return out[:3], out[3:]

    def batch(self, offsets, precision=SHP_PREC_FP32C):
        """Plan for a batch of independent plumes held in one concatenated
        DeviceState; instance k owns ions [offsets[k], offsets[k+1])."""
        return BatchPlan(self, offsets, precision)


class BatchPlan:
    """Persistent-launch direct sum over many plume instances (C ABI shp_batch_*)."""

    N_CONFIGS = 4

    def __init__(self, eng, offsets, precision):
        self.eng = eng
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        self.ninst = off.shape[0] - 1
        self.n = int(off[-1])
        h = ctypes.c_void_p()
        _check(lib.shp_batch_create(eng.h, self.ninst, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                    precision, ctypes.byref(h)), "shp_batch_create")
        self.h = h

    def close(self):
        if self.h:
            lib.shp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tune(self, state):
        """Times the tile-shape candidates; returns (ms per candidate, chosen index)."""
        ms = (ctypes.c_double * self.N_CONFIGS)()
        ch = ctypes.c_int()
        p = state.as_struct()
        _torch_ready(state.x.device)
        _check(lib.shp_batch_tune(self.h, ctypes.byref(p), ms, ctypes.byref(ch)), "shp_batch_tune")
        return list(ms), ch.value

    def set_config(self, k):
        _check(lib.shp_batch_set_config(self.h, int(k)), "shp_batch_set_config")

    def forces(self, state, out=None):
        import torch
        if out is None:
            out = tuple(torch.empty(self.n, dtype=torch.float64, device=state.x.device)
                        for _ in range(3))
        p = state.as_struct()
        _torch_ready(state.x.device)
        _check(lib.shp_batch_forces(self.h, ctypes.byref(p), out[0].data_ptr(), out[1].data_ptr(),
                                    out[2].data_ptr()), "shp_batch_forces")
        self.eng.sync()
        return out

    def step_vjp(self, state, dt, xb1, vb1, qb):
        """VJP of one DKD step of every instance; xb1, vb1 are (3, n) float64
        device tensors (component-major); returns (xb0, vb0) and adds into qb."""
        import torch
        xb0 = torch.empty_like(xb1)
        vb0 = torch.empty_like(vb1)
        s = state.as_state()
        _torch_ready(state.x.device)
        _check(lib.shp_batch_step_vjp(self.h, ctypes.byref(s), float(dt), xb1.data_ptr(),
                                      vb1.data_ptr(), xb0.data_ptr(), vb0.data_ptr(), qb.data_ptr()),
               "shp_batch_step_vjp")
        self.eng.sync()
        return xb0, vb0

    def leapfrog(self, state, opts, t0, dt, nsteps, use_graph=True, acc=None):
        import torch
        if acc is None:
            acc = tuple(torch.zeros(self.n, dtype=torch.float64, device=state.x.device)
                        for _ in range(3))
        s = state.as_state()
        _torch_ready(state.x.device)
        _check(lib.shp_batch_leapfrog(self.h, ctypes.byref(s), ctypes.byref(opts), float(t0),
                                      float(dt), int(nsteps), 1 if use_graph else 0,
                                      acc[0].data_ptr(), acc[1].data_ptr(), acc[2].data_ptr()),
               "shp_batch_leapfrog")
        self.eng.sync()
        return acc


def make_step_opts(method=0, precision=2, symmetry=None, use_efield=False, bc_space=(0, 0, 0, 0),
                   bc_time=0.0, fmm_order=10, leaf_size=20, theta=0.5, fmm_ring=2, leaf_level=0):
    o = StepOpts()
    o.method = method
    o.precision = precision
    sym = symmetry if symmetry is not None else make_symmetry("none")
    o.sym = sym
    o._sym = sym  # keeps the lattice array alive
    o.use_efield = 1 if use_efield else 0
    for c in range(4):
        o.bc_space[c] = float(bc_space[c])
    o.bc_time = float(bc_time)
    o.fmm_order = int(fmm_order)
    o.leaf_size = int(leaf_size)
    o.theta = float(theta)
    o.fmm_ring = int(fmm_ring)
    o.leaf_level = int(leaf_level)
    return o


def _engine_efield_setup(self, mesh):
    """Upload a tip-field mesh dict (xy, er, ez, tri, nbr, r0, E0) and build the locator."""
    keep = {k: np.ascontiguousarray(mesh[k], dtype=np.float64) for k in ("xy", "er", "ez")}
    keep.update({k: np.ascontiguousarray(mesh[k], dtype=np.int32) for k in ("tri", "nbr")})
    m = FieldMesh()
    m.nnodes = keep["er"].shape[0]
    m.xy, m.er, m.ez = keep["xy"].ctypes.data, keep["er"].ctypes.data, keep["ez"].ctypes.data
    m.ntri = keep["tri"].shape[0]
    m.tri, m.nbr = keep["tri"].ctypes.data, keep["nbr"].ctypes.data
    m.r0, m.E0 = float(mesh["r0"]), float(mesh["E0"])
    _check(lib.shp_efield_setup(self.h, ctypes.byref(m)), "shp_efield_setup")


def _engine_efield_dev(self, x, y, z, want_tri=False):
    import torch
    n = int(x.shape[0])
    out = tuple(torch.empty(n, dtype=torch.float64, device=x.device) for _ in range(3))
    tri = torch.empty(n, dtype=torch.int32, device=x.device) if want_tri else None
    _torch_ready(x.device)
    _check(lib.shp_efield_dev(self.h, n, x.data_ptr(), y.data_ptr(), z.data_ptr(),
                              out[0].data_ptr(), out[1].data_ptr(), out[2].data_ptr(),
                              tri.data_ptr() if tri is not None else None), "shp_efield_dev")
    self.sync()
    return (out, tri) if want_tri else out


def _engine_leapfrog_run(self, state, opts, t0, dt, nsteps, acc=None):
    import torch
    n = state.n
    if acc is None:
        acc = tuple(torch.zeros(n, dtype=torch.float64, device=state.x.device) for _ in range(3))
    s = state.as_state()
    _torch_ready(state.x.device)
    _check(lib.shp_leapfrog_run(self.h, ctypes.byref(s), ctypes.byref(opts), float(t0), float(dt),
                                int(nsteps), acc[0].data_ptr(), acc[1].data_ptr(),
                                acc[2].data_ptr()), "shp_leapfrog_run")
    self.sync()
    return acc


def _engine_leapfrog_substeps(self, state, opts, levels, level, t0, dt, nsteps, sub=None, acc=None):
    """nsteps of the multi-rate LeapfrogDKDSubsteps on a DeviceState; levels as for
    make_substep_levels; level: int32 device tensor of the ions' levels (in/out)."""
    import torch
    n = state.n
    if acc is None:
        acc = tuple(torch.zeros(n, dtype=torch.float64, device=state.x.device) for _ in range(3))
    s = state.as_state()
    table = make_substep_levels(levels)
    _torch_ready(state.x.device)
    _check(lib.shp_leapfrog_substeps_run(self.h, ctypes.byref(s), ctypes.byref(opts),
                                         ctypes.byref(sub) if sub is not None else None,
                                         ctypes.addressof(table), len(levels) - 1, level.data_ptr(),
                                         float(t0), float(dt), int(nsteps), acc[0].data_ptr(),
                                         acc[1].data_ptr(), acc[2].data_ptr()), "shp_leapfrog_substeps_run")
    self.sync()
    return acc


def _engine_fmm_dev(self, state, order=8, leaf_level=0, leaf_size=32, ring=2, out=None,
                    symmetry=None):
    """FMM accelerations on a DeviceState (with symmetry images when given);
    returns (ax, ay, az) float64 device tensors."""
    import torch
    n = state.n
    if out is None:
        out = tuple(torch.empty(n, dtype=torch.float64, device=state.x.device) for _ in range(3))
    p = state.as_struct()
    _torch_ready(state.x.device)
    if symmetry is None:
        _check(lib.shp_fmm_dev(self.h, ctypes.byref(p), int(order), int(leaf_level),
                               int(leaf_size), int(ring), out[0].data_ptr(), out[1].data_ptr(),
                               out[2].data_ptr()), "shp_fmm_dev")
    else:
        _check(lib.shp_fmm_sym_dev(self.h, ctypes.byref(p), ctypes.byref(symmetry), int(order),
                                   int(leaf_level), int(leaf_size), int(ring), out[0].data_ptr(),
                                   out[1].data_ptr(), out[2].data_ptr()), "shp_fmm_sym_dev")
    self.sync()
    return out


def _engine_fmm_set_type(self, fmm_type):
    """FmmType of later FMM evaluations (SHP_FMM_TREECODE, _TREECODE_OUTLIERS,
    _FMM (default), _FMM_OUTLIERS)."""
    _check(lib.shp_fmm_set_type(self.h, int(fmm_type)), "shp_fmm_set_type")


def _engine_fmm_outliers(self, state):
    """Outlier flags (int32 numpy array) of the outlier modes, and their number."""
    import torch
    flags = torch.zeros(state.n, dtype=torch.int32, device=state.x.device)
    p = state.as_struct()
    _torch_ready(state.x.device)
    nout = ctypes.c_int(0)
    _check(lib.shp_fmm_outliers_dev(self.h, ctypes.byref(p), flags.data_ptr(), ctypes.byref(nout)),
           "shp_fmm_outliers_dev")
    self.sync()
    return flags.cpu().numpy(), nout.value


def _engine_fmm_leaf_level(self):
    return lib.shp_fmm_last_leaf_level(self.h)


FMM_STAT_KEYS = ("p2m_m2m_ms", "ttable_ms", "m2l_ms", "l2l_ms", "leaf_ms", "m2l_pairs",
                 "p2p_pairs", "leaf_level", "m2l_madds", "m2l_pass0_ms", "m2l_pass0_madds")


def _engine_fmm_stats(self):
    """Measurements of the last FMM evaluation (dict of FMM_STAT_KEYS)."""
    out = (ctypes.c_double * len(FMM_STAT_KEYS))()
    _check(lib.shp_fmm_stats(self.h, out), "shp_fmm_stats")
    return dict(zip(FMM_STAT_KEYS, list(out)))


def _engine_fmm_timing(self, reset=False):
    """Sums of the per-call FMM measurements since the last reset, plus "calls"."""
    out = (ctypes.c_double * (len(FMM_STAT_KEYS) + 1))()
    _check(lib.shp_fmm_timing(self.h, out, 1 if reset else 0), "shp_fmm_timing")
    d = dict(zip(FMM_STAT_KEYS, list(out)))
    d["calls"] = int(out[len(FMM_STAT_KEYS)])
    return d


def _engine_fmm_level_estimates(self):
    """The last FMM call's leaf-level candidates: list of dicts (level,
    p2p_pairs, far_pairs, est_ms); empty when the level was not chosen."""
    out = (ctypes.c_double * (4 * 24))()
    n = lib.shp_fmm_level_estimates(self.h, out, 24)
    _check(min(n, 0), "shp_fmm_level_estimates")
    return [dict(level=int(out[4 * k]), p2p_pairs=out[4 * k + 1], far_pairs=out[4 * k + 2],
                 est_ms=out[4 * k + 3]) for k in range(min(n, 24))]


def _engine_compare_accel(self, status, a, ref):
    """Accuracy of a = (ax, ay, az) against ref (device float64 tensors) over
    the ions whose int32 status is 1, as the reference's compareAcceleration
    defines it: dict(n, l2, max, idx)."""
    out = (ctypes.c_double * 4)()
    n = status.shape[0]
    for t in tuple(a) + tuple(ref):
        assert t.shape[0] == n and t.is_contiguous()
    _torch_ready(status.device)
    _check(lib.shp_compare_accel(self.h, n, status.data_ptr(), a[0].data_ptr(), a[1].data_ptr(),
                                 a[2].data_ptr(), ref[0].data_ptr(), ref[1].data_ptr(),
                                 ref[2].data_ptr(), out), "shp_compare_accel")
    return dict(n=int(out[0]), l2=out[1], max=out[2], idx=int(out[3]))


Engine.fmm_dev = _engine_fmm_dev
Engine.fmm_level_estimates = _engine_fmm_level_estimates
Engine.compare_accel = _engine_compare_accel
Engine.fmm_leaf_level = _engine_fmm_leaf_level
Engine.fmm_set_type = _engine_fmm_set_type
Engine.fmm_outliers = _engine_fmm_outliers
Engine.leapfrog_substeps = _engine_leapfrog_substeps
Engine.fmm_stats = _engine_fmm_stats
Engine.fmm_timing = _engine_fmm_timing


def _engine_last_update_ms(self):
    return lib.shp_last_update_ms(self.h)


Engine.efield_setup = _engine_efield_setup
Engine.efield_dev = _engine_efield_dev
Engine.leapfrog_run = _engine_leapfrog_run
Engine.last_update_ms = _engine_last_update_ms


class DeviceState:
    """FP64 SoA particle state resident in HBM (torch tensors as plain storage)."""

    def __init__(self, x, y, z, q, status, species, vx=None, vy=None, vz=None, t_inj=None):
        self.x, self.y, self.z, self.q = x, y, z, q
        self.status, self.species = status, species
        self.vx, self.vy, self.vz, self.t_inj = vx, vy, vz, t_inj
        self.n = int(x.shape[0])

    @classmethod
    def from_numpy(cls, d, device="cuda:0"):
        import torch

        def f64(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)

        def i32(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)

        kw = {}
        for k in ("vx", "vy", "vz", "t_inj"):
            if k in d:
                kw[k] = f64(d[k])
        return cls(f64(d["x"]), f64(d["y"]), f64(d["z"]), f64(d["q"]), i32(d["status"]),
                   i32(d["species"]), **kw)

    def as_state(self):
        s = State()
        s.n = self.n
        for k in ("x", "y", "z", "vx", "vy", "vz", "q", "status", "species", "t_inj"):
            setattr(s, k, getattr(self, k).data_ptr())
        return s

    def to_numpy(self):
        out = {}
        for k in ("x", "y", "z", "vx", "vy", "vz", "q", "status", "species", "t_inj"):
            v = getattr(self, k)
            if v is not None:
                out[k] = v.cpu().numpy()
        return out

    def as_struct(self):
        p = Particles()
        p.n = self.n
        p.x, p.y, p.z = self.x.data_ptr(), self.y.data_ptr(), self.z.data_ptr()
        p.q = self.q.data_ptr()
        p.status = self.status.data_ptr()
        p.species = self.species.data_ptr()
        return p
