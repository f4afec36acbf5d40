// Direct softened-Coulomb P2P sums on gfx950.
//
// One workgroup owns a tile of targets (one or two per lane) and a slice of the
// source list ("split-j"); sources are staged through LDS 256 at a time and read
// back as wave-wide broadcasts.  Partial sums per split land in a workspace and
// a second kernel combines them in a fixed order, so results are deterministic.
//
// Arithmetic contracts (see include/solveit_hip.h):
//   REF   float displacements, double everything else, as
//         DirectForceFunction::direct (direct_force_method.cpp:153-190) and its
//         rotational/translational siblings;
//   FP64  double positions and displacements;
//   FP32C positions re-expressed relative to a per-workgroup origin (double ->
//         float once per staged source), float pair arithmetic with v_rsq_f32,
//         float partial sums per 256-source tile folded into a per-target
//         two-sum (hi, lo) accumulator.  Targets are visited in Morton order so
//         a workgroup's targets are spatially compact and the origin shift keeps
//         near-pair displacements accurate to float rounding.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <mutex>
#include <vector>

#include "shp_internal.h"
#include "direct_kernels.h"

namespace shp {

static inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// Device buffers of one evaluation: the handle's grow-only workspace, or, for
// the allocation-free entry point, consecutive 256-byte aligned pieces of the
// per-device scratch array (base != nullptr).
struct Scratch {
    shp_handle h;
    char* base = nullptr;
    size_t cap = 0, used = 0;
    void* get(int slot, size_t bytes) {
        if (!base) return h->get(slot, bytes);
        const size_t off = (used + 255) & ~(size_t)255;
        if (off + bytes > cap) {
            set_error("shp_direct_dev_noalloc: the evaluation needs more than the %zu bytes of scratch", cap);
            return nullptr;
        }
        used = off + bytes;
        return base + off;
    }
};

#define SC_GET(ptr, type, sc, slot, count)                                                \
    type* ptr = static_cast<type*>((sc).get((slot), sizeof(type) * (size_t)(count)));     \
    if (!ptr) return (sc).base ? SHP_ERR_INVALID : SHP_ERR_NOMEM

// Most images the allocation-free entry point takes (rotational sectors or
// lattice points).
constexpr int kNoAllocMaxImages = 1024;

// Per-image transforms: ROT (c, s), TRANS (ox, oy, oz).  For the REF contract
// the translational offset is the float product lattice * (float)domain, as
// direct_force_method_translational.cpp:324-327 computes it; rotations use
// float cos/sin (exact integers for 4 sectors, auxiliary_functions.h:55-57).
template <typename Vec>
static int build_images_into(shp_handle h, Scratch& sc, const shp_symmetry* sym, bool ref, Vec& im,
                             const double4** out, int* nimg) {
    if (!sym || sym->type == SHP_SYM_NONE) {
        im.push_back(make_double4(1.0, 0.0, 0.0, 0.0));
    } else if (sym->type == SHP_SYM_ROTATIONAL) {
        SHP_REQUIRE(sym->nsec >= 1 && sym->nsec <= 4096, "rotational symmetry: nsec=%d", sym->nsec);
        for (int k = 0; k < sym->nsec; k++) {
            double c, s;
            if ((4 * k) % sym->nsec == 0) {
                const int q = (4 * k / sym->nsec) & 3;
                const int cs[4] = {1, 0, -1, 0};
                c = cs[q];
                s = cs[(q + 3) & 3];
            } else {
                const double ang = 2.0 * 3.14159265358979323846 * k / sym->nsec;
                c = cos(ang);
                s = sin(ang);
                if (ref) {
                    c = (double)(float)c;
                    s = (double)(float)s;
                }
            }
            im.push_back(make_double4(c, s, 0.0, 0.0));
        }
    } else if (sym->type == SHP_SYM_TRANSLATIONAL) {
        SHP_REQUIRE(sym->nimg >= 1 && sym->lattice != nullptr,
                    "translational symmetry needs a lattice (nimg=%d)", sym->nimg);
        for (int k = 0; k < sym->nimg; k++) {
            double o[3];
            for (int c = 0; c < 3; c++) {
                const int l = sym->lattice[3 * k + c];
                o[c] = ref ? (double)((float)l * (float)sym->domain[c]) : l * sym->domain[c];
            }
            im.push_back(make_double4(o[0], o[1], o[2], 0.0));
        }
    } else {
        SHP_REQUIRE(false, "unknown symmetry type %d", sym->type);
    }
    SC_GET(d, double4, sc, WS_IMAGES, im.size());
    SHP_HIP_CHECK(hipMemcpyAsync(d, im.data(), im.size() * sizeof(double4), hipMemcpyHostToDevice,
                                 h->stream));
    // The host copy dies on return: wait for the copy to leave it.
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    *out = d;
    *nimg = (int)im.size();
    return SHP_OK;
}

// A fixed-capacity host list of images (no heap allocation), for the
// allocation-free entry point.
struct ImageList {
    double4 v[kNoAllocMaxImages];
    size_t n = 0;
    void push_back(const double4& x) {
        if (n < kNoAllocMaxImages) v[n] = x;
        n++;
    }
    size_t size() const { return n; }
    const double4* data() const { return v; }
};

int build_images(shp_handle h, const shp_symmetry* sym, bool ref, const double4** out, int* nimg,
                 Scratch* scp = nullptr) {
    Scratch ws{h};
    Scratch& sc = scp ? *scp : ws;
    if (sc.base) {
        const int want = !sym || sym->type == SHP_SYM_NONE
                             ? 1
                             : (sym->type == SHP_SYM_ROTATIONAL ? sym->nsec : sym->nimg);
        SHP_REQUIRE(want <= kNoAllocMaxImages, "shp_direct_dev_noalloc: %d images (at most %d)", want,
                    kNoAllocMaxImages);
        ImageList im;  // on the stack: no heap allocation
        return build_images_into(h, sc, sym, ref, im, out, nimg);
    }
    std::vector<double4> im;
    return build_images_into(h, sc, sym, ref, im, out, nimg);
}

static int choose_split(shp_handle h, int target_blocks, int nsrc_upper) {
    const int want = h->num_cus * 8;
    int ns = cdiv(want, std::max(1, target_blocks));
    ns = std::min(ns, std::max(1, nsrc_upper / 1024));
    return std::max(1, std::min(ns, 256));
}

// Compact [j0,j1) sources into device records; returns device count pointer.
static int stage_sources_f64(shp_handle h, Scratch& sc, const double* X, const double* Y,
                             const double* Z, const double* Q, const int* status, int j0, int j1,
                             const double4** src, const int** count) {
    const int n = j1 - j0;
    SC_GET(flags, char, sc, WS_FLAGS, n);
    SC_GET(idx, int, sc, WS_KEYS, n);
    SC_GET(cnt, int, sc, WS_COUNTS, 4);
    SC_GET(rec, double4, sc, WS_SRC, n);
    hipLaunchKernelGGL(k_flag_sources_f64, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, status, j0,
                       n, flags);
    size_t tmp = 0;
    hipcub::CountingInputIterator<int> it(j0);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tmp, it, flags, idx, cnt, n, h->stream));
    SC_GET(cub, char, sc, WS_CUB, tmp);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, tmp, it, flags, idx, cnt, n, h->stream));
    hipLaunchKernelGGL(k_gather_sources_f64, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, idx, cnt,
                       X, Y, Z, Q, n, rec);
    *src = rec;
    *count = cnt;
    return SHP_OK;
}

static int stage_sources_ref(shp_handle h, const float4* pos4, const float4* acc4, int j0, int j1,
                             const float4** src, const int** count) {
    const int n = j1 - j0;
    SHP_WS(flags, char, h, WS_FLAGS, n);
    SHP_WS(idx, int, h, WS_KEYS, n);
    SHP_WS(cnt, int, h, WS_COUNTS, 4);
    SHP_WS(rec, float4, h, WS_SRC, n);
    hipLaunchKernelGGL(k_flag_sources_ref, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, acc4, j0, n,
                       flags);
    size_t tmp = 0;
    hipcub::CountingInputIterator<int> it(j0);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tmp, it, flags, idx, cnt, n, h->stream));
    SHP_WS(cub, char, h, WS_CUB, tmp);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, tmp, it, flags, idx, cnt, n, h->stream));
    hipLaunchKernelGGL(k_gather_sources_ref, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, idx, cnt,
                       pos4, n, rec);
    *src = rec;
    *count = cnt;
    return SHP_OK;
}

// Target slots in Morton order (status-1 particles first, others last).
static int morton_targets(shp_handle h, Scratch& sc, const double* X, const double* Y,
                          const double* Z, const int* status, int i0, int i1, const int** tgt) {
    const int n = i1 - i0;
    const int nb = std::min(cdiv(n, 256), 1024);
    SC_GET(bpart, double, sc, WS_BBOX, nb * 6);
    SC_GET(keys, uint32_t, sc, WS_KEYS, n);
    SC_GET(vals, int, sc, WS_TGT, n);
    SC_GET(keys2, uint32_t, sc, WS_KEYS2, n);
    SC_GET(vals2, int, sc, WS_VALS2, n);
    SC_GET(box, double, sc, WS_BOX, 8);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(kThreads), 0, h->stream, X, Y, Z, status, i0,
                       n, bpart);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(64), 0, h->stream, bpart, nb, box);
    hipLaunchKernelGGL(k_morton30, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, X, Y, Z, status, i0,
                       n, box, keys, vals);
    size_t tmp = 0;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, keys2, vals, vals2, n, 0, 32,
                                                     h->stream));
    SC_GET(cub, char, sc, WS_CUB, tmp);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub, tmp, keys, keys2, vals, vals2, n, 0, 32,
                                                     h->stream));
    *tgt = vals2;
    return SHP_OK;
}

// Shared FP64-state driver: arrays are device SoA; targets [i0,i1), sources
// [j0,j1).  mode 0 writes ax/ay/az[i] (zeroed first over [0, nout)); mode 1
// writes out3 rows (i - i0).
static int direct_f64_state(shp_handle h, int prec, const double* X, const double* Y,
                            const double* Z, const double* Q, const int* status, const int* species,
                            int i0, int i1, int j0, int j1, const shp_symmetry* sym, int mode,
                            int nout, double* ax, double* ay, double* az, double* out3,
                            Scratch* scp = nullptr) {
    Scratch ws{h};
    Scratch& sc = scp ? *scp : ws;
    const int nslots = i1 - i0;
    const double4* img;
    int nimg;
    int rc = build_images(h, sym, false, &img, &nimg, &sc);
    if (rc) return rc;
    const double4* src;
    const int* nsrc;
    rc = stage_sources_f64(h, sc, X, Y, Z, Q, status, j0, j1, &src, &nsrc);
    if (rc) return rc;
    const int* tgt;
    if (prec == SHP_PREC_FP32C) {
        rc = morton_targets(h, sc, X, Y, Z, status, i0, i1, &tgt);
        if (rc) return rc;
    } else {
        SC_GET(ident, int, sc, WS_TGT, nslots);
        hipLaunchKernelGGL(k_iota, dim3(cdiv(nslots, 256)), dim3(256), 0, h->stream, ident, i0,
                           nslots);
        tgt = ident;
    }
    const int per_block = prec == SHP_PREC_FP32C ? kThreads * kF32TargetsPerLane : kThreads;
    const int tb = cdiv(nslots, per_block);
    const int nsplit = choose_split(h, tb, j1 - j0);
    SC_GET(part, double, sc, WS_PARTIAL, (size_t)nsplit * 3 * nslots);
    if (mode == 0) {
        hipLaunchKernelGGL(k_fill, dim3(cdiv(nout, 256)), dim3(256), 0, h->stream, ax, (size_t)nout, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(cdiv(nout, 256)), dim3(256), 0, h->stream, ay, (size_t)nout, 0.0);
        hipLaunchKernelGGL(k_fill, dim3(cdiv(nout, 256)), dim3(256), 0, h->stream, az, (size_t)nout, 0.0);
    } else {
        hipLaunchKernelGGL(k_fill, dim3(cdiv(3 * (size_t)nslots, 256)), dim3(256), 0, h->stream, out3,
                           3 * (size_t)nslots, 0.0);
    }
    const int symt = sym ? sym->type : SHP_SYM_NONE;
    dim3 grid(tb, nsplit);
    SHP_HIP_CHECK(hipEventRecord(h->ev0, h->stream));
    if (prec == SHP_PREC_FP32C) {
        const float eps = (float)kSoftening;
        if (symt == SHP_SYM_NONE)
            hipLaunchKernelGGL(k_direct_f32c<SHP_SYM_NONE>, grid, dim3(kThreads), 0, h->stream, tgt,
                               nslots, X, Y, Z, src, nsrc, img, nimg, nsplit, eps, part, nslots);
        else if (symt == SHP_SYM_ROTATIONAL)
            hipLaunchKernelGGL(k_direct_f32c<SHP_SYM_ROTATIONAL>, grid, dim3(kThreads), 0, h->stream,
                               tgt, nslots, X, Y, Z, src, nsrc, img, nimg, nsplit, eps, part, nslots);
        else
            hipLaunchKernelGGL(k_direct_f32c<SHP_SYM_TRANSLATIONAL>, grid, dim3(kThreads), 0,
                               h->stream, tgt, nslots, X, Y, Z, src, nsrc, img, nimg, nsplit, eps,
                               part, nslots);
    } else {
        if (symt == SHP_SYM_NONE)
            hipLaunchKernelGGL(k_direct_f64<SHP_SYM_NONE>, grid, dim3(kThreads), 0, h->stream, tgt,
                               nslots, X, Y, Z, src, nsrc, img, nimg, nsplit, kSoftening, part, nslots);
        else if (symt == SHP_SYM_ROTATIONAL)
            hipLaunchKernelGGL(k_direct_f64<SHP_SYM_ROTATIONAL>, grid, dim3(kThreads), 0, h->stream,
                               tgt, nslots, X, Y, Z, src, nsrc, img, nimg, nsplit, kSoftening, part,
                               nslots);
        else
            hipLaunchKernelGGL(k_direct_f64<SHP_SYM_TRANSLATIONAL>, grid, dim3(kThreads), 0,
                               h->stream, tgt, nslots, X, Y, Z, src, nsrc, img, nimg, nsplit,
                               kSoftening, part, nslots);
    }
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipEventRecord(h->ev1, h->stream));
    h->timed = true;
    hipLaunchKernelGGL(k_reduce_scatter<false>, dim3(cdiv(nslots, 256)), dim3(256), 0, h->stream, tgt,
                       nslots, nsplit, part, nslots, status, Q, species, nullptr, nullptr, nullptr,
                       k_const_host(), mode, i0, ax, ay, az, out3);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

// out3 rows (ax, ay, az) of the targets (status 1), zero rows for the others.
__global__ void k_soa_to_rows(const double* __restrict__ ax, const double* __restrict__ ay,
                              const double* __restrict__ az, const int* __restrict__ status, int m,
                              double* __restrict__ out3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const bool t = status[i] == 1;
    out3[3 * (size_t)i] = t ? ax[i] : 0.0;
    out3[3 * (size_t)i + 1] = t ? ay[i] : 0.0;
    out3[3 * (size_t)i + 2] = t ? az[i] : 0.0;
}

// ------------------------------------------------ allocation-free variant ----

// Scratch of shp_direct_dev_noalloc: one copy per GPU the library runs on,
// held until the process ends.  Host code never uses this name as an address:
// the per-device address comes from hipGetSymbolAddress at set-up.
static __device__ __align__(256) char scratch[SHP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch_ptr[kMaxDevices] = {};

// Looks up this device's copy of the scratch array (the device must be
// current); called once per device from shp_create.
int direct_scratch_setup(int device) {
    SHP_REQUIRE(device >= 0 && device < kMaxDevices, "scratch set-up: device %d", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return SHP_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr) {
        set_error("scratch set-up on device %d: hipGetSymbolAddress failed: %s", device, hipGetErrorString(e));
        return SHP_ERR_HIP;
    }
    g_scratch_ptr[device] = (char*)ptr;
    return SHP_OK;
}

static char* scratch_of(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace shp

using namespace shp;

extern "C" {

int shp_direct_dev_noalloc(shp_handle h, const shp_particles* p, int precision, const shp_symmetry* sym,
                           double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && p && ax && ay && az, "shp_direct_dev_noalloc: null argument");
    SHP_REQUIRE(p->n >= 0 && p->n <= SHP_NOALLOC_MAX_N, "shp_direct_dev_noalloc: n=%d (at most %d)", p->n,
                SHP_NOALLOC_MAX_N);
    SHP_REQUIRE(precision == SHP_PREC_FP64 || precision == SHP_PREC_FP32C,
                "shp_direct_dev_noalloc: precision %d not supported on FP64 state", precision);
    if (p->n == 0) return SHP_OK;
    SHP_REQUIRE(p->x && p->y && p->z && p->q && p->status && p->species,
                "shp_direct_dev_noalloc: null particle array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    Scratch sc{h};
    sc.base = scratch_of(h->device);
    sc.cap = SHP_NOALLOC_SCRATCH_BYTES;
    if (!sc.base) {
        set_error("shp_direct_dev_noalloc: no scratch on device %d (handle not set up)", h->device);
        return SHP_ERR_HIP;
    }
    return direct_f64_state(h, precision, p->x, p->y, p->z, p->q, p->status, p->species, 0, p->n, 0,
                            p->n, sym, 0, p->n, ax, ay, az, nullptr, &sc);
}

int shp_direct_dev(shp_handle h, const shp_particles* p, int precision, const shp_symmetry* sym,
                   double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && p && ax && ay && az, "shp_direct_dev: null argument");
    SHP_REQUIRE(p->n >= 0, "shp_direct_dev: n=%d", p->n);
    SHP_REQUIRE(precision == SHP_PREC_FP64 || precision == SHP_PREC_FP32C,
                "shp_direct_dev: precision %d not supported on FP64 state", precision);
    if (p->n == 0) return SHP_OK;
    SHP_REQUIRE(p->x && p->y && p->z && p->q && p->status && p->species,
                "shp_direct_dev: null particle array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    return direct_f64_state(h, precision, p->x, p->y, p->z, p->q, p->status, p->species, 0, p->n, 0,
                            p->n, sym, 0, p->n, ax, ay, az, nullptr);
}

int shp_direct_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n,
                    int i0, int i1, int j0, int j1, int precision, const shp_symmetry* sym,
                    double* out3) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4 && out3, "shp_direct_host: null argument");
    SHP_REQUIRE(n >= 0 && 0 <= i0 && i0 <= i1 && i1 <= n && 0 <= j0 && j0 <= j1 && j1 <= n,
                "shp_direct_host: bad ranges n=%d [%d,%d) [%d,%d)", n, i0, i1, j0, j1);
    SHP_REQUIRE(precision == SHP_PREC_REF || precision == SHP_PREC_FP32C,
                "shp_direct_host: precision %d not supported on float AoS input", precision);
    const int nt = i1 - i0;
    if (nt == 0) return SHP_OK;
    if (j1 == j0) {
        for (int k = 0; k < 3 * nt; k++) out3[k] = 0.0;
        return SHP_OK;
    }
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(dpos, float4, h, WS_HOST_IN, n);
    SHP_WS(dvel, float4, h, WS_HOST_IN2, n);
    SHP_WS(dacc, float4, h, WS_HOST_IN3, n);
    SHP_WS(dout, double, h, WS_HOST_OUT, 3 * (size_t)nt);
    SHP_HIP_CHECK(hipMemcpyAsync(dpos, pos4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dvel, vel4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dacc, acc4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    int rc;
    if (precision == SHP_PREC_FP32C) {
        // Widen the float AoS state to FP64 SoA and use the FP64-state driver.
        SHP_WS(X, double, h, WS_WIDE_F64, 4 * (size_t)n);
        SHP_WS(ints, int, h, WS_WIDE_I32, 2 * (size_t)n);
        double *Y = X + n, *Z = Y + n, *Q = Z + n;
        int *st = ints, *sp = ints + n;
        hipLaunchKernelGGL(k_aos_to_f64, dim3(cdiv(n, 256)), dim3(256), 0, h->stream, dpos, dvel,
                           dacc, n, X, Y, Z, Q, st, sp);
        rc = direct_f64_state(h, SHP_PREC_FP32C, X, Y, Z, Q, st, sp, i0, i1, j0, j1, sym, 1, 0,
                              nullptr, nullptr, nullptr, dout);
        if (rc) return rc;
    } else {
        const double4* img;
        int nimg;
        rc = build_images(h, sym, true, &img, &nimg);
        if (rc) return rc;
        const float4* src;
        const int* nsrc;
        rc = stage_sources_ref(h, dpos, dacc, j0, j1, &src, &nsrc);
        if (rc) return rc;
        SHP_WS(ident, int, h, WS_TGT, nt);
        hipLaunchKernelGGL(k_iota, dim3(cdiv(nt, 256)), dim3(256), 0, h->stream, ident, i0, nt);
        const int tb = cdiv(nt, kThreads);
        const int nsplit = choose_split(h, tb, j1 - j0);
        SHP_WS(part, double, h, WS_PARTIAL, (size_t)nsplit * 3 * nt);
        hipLaunchKernelGGL(k_fill, dim3(cdiv(3 * (size_t)nt, 256)), dim3(256), 0, h->stream, dout,
                           3 * (size_t)nt, 0.0);
        const int symt = sym ? sym->type : SHP_SYM_NONE;
        dim3 grid(tb, nsplit);
        SHP_HIP_CHECK(hipEventRecord(h->ev0, h->stream));
        if (symt == SHP_SYM_NONE)
            hipLaunchKernelGGL(k_direct_ref<SHP_SYM_NONE>, grid, dim3(kThreads), 0, h->stream, ident,
                               nt, dpos, src, nsrc, img, nimg, nsplit, kSoftening, part, nt);
        else if (symt == SHP_SYM_ROTATIONAL)
            hipLaunchKernelGGL(k_direct_ref<SHP_SYM_ROTATIONAL>, grid, dim3(kThreads), 0, h->stream,
                               ident, nt, dpos, src, nsrc, img, nimg, nsplit, kSoftening, part, nt);
        else
            hipLaunchKernelGGL(k_direct_ref<SHP_SYM_TRANSLATIONAL>, grid, dim3(kThreads), 0,
                               h->stream, ident, nt, dpos, src, nsrc, img, nimg, nsplit, kSoftening,
                               part, nt);
        SHP_HIP_CHECK(hipGetLastError());
        SHP_HIP_CHECK(hipEventRecord(h->ev1, h->stream));
        h->timed = true;
        hipLaunchKernelGGL(k_reduce_scatter<true>, dim3(cdiv(nt, 256)), dim3(256), 0, h->stream,
                           ident, nt, nsplit, part, nt, nullptr, nullptr, nullptr, dpos, dvel, dacc,
                           k_const_host(), 1, i0, nullptr, nullptr, nullptr, dout);
        SHP_HIP_CHECK(hipGetLastError());
    }
    SHP_HIP_CHECK(hipMemcpyAsync(out3, dout, sizeof(double) * 3 * (size_t)nt, hipMemcpyDeviceToHost,
                                 h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

// FMM on the reference's host AoS arrays: the ions [i0, i1) are widened to
// FP64 SoA on the device and evaluated by shp_fmm_sym_dev.
int shp_fmm_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n, int i0,
                 int i1, int order, int leaf_level, int leaf_size, int ring, const shp_symmetry* sym,
                 double* out3) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4 && out3, "shp_fmm_host: null argument");
    SHP_REQUIRE(n >= 0 && 0 <= i0 && i0 <= i1 && i1 <= n, "shp_fmm_host: bad range n=%d [%d,%d)", n, i0, i1);
    const int m = i1 - i0;
    if (m == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(dpos, float4, h, WS_HOST_IN, m);
    SHP_WS(dvel, float4, h, WS_HOST_IN2, m);
    SHP_WS(dacc, float4, h, WS_HOST_IN3, m);
    SHP_WS(X, double, h, WS_WIDE_F64, 7 * (size_t)m);
    SHP_WS(ints, int, h, WS_WIDE_I32, 2 * (size_t)m);
    double *Y = X + m, *Z = Y + m, *Q = Z + m, *AX = Q + m, *AY = AX + m, *AZ = AY + m;
    int *st = ints, *sp = ints + m;
    SHP_WS(dout, double, h, WS_HOST_OUT, 3 * (size_t)m);
    SHP_HIP_CHECK(hipMemcpyAsync(dpos, pos4 + 4 * (size_t)i0, sizeof(float4) * m, hipMemcpyHostToDevice,
                                 h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dvel, vel4 + 4 * (size_t)i0, sizeof(float4) * m, hipMemcpyHostToDevice,
                                 h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dacc, acc4 + 4 * (size_t)i0, sizeof(float4) * m, hipMemcpyHostToDevice,
                                 h->stream));
    hipLaunchKernelGGL(k_aos_to_f64, dim3(cdiv(m, 256)), dim3(256), 0, h->stream, dpos, dvel, dacc, m, X, Y,
                       Z, Q, st, sp);
    shp_particles p;
    p.n = m;
    p.x = X;
    p.y = Y;
    p.z = Z;
    p.q = Q;
    p.status = st;
    p.species = sp;
    const int rc = shp_fmm_sym_dev(h, &p, sym, order, leaf_level, leaf_size, ring, AX, AY, AZ);
    if (rc) return rc;
    hipLaunchKernelGGL(k_soa_to_rows, dim3(cdiv(m, 256)), dim3(256), 0, h->stream, AX, AY, AZ, st, m, dout);
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipMemcpyAsync(out3, dout, sizeof(double) * 3 * (size_t)m, hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

}  // extern "C"
