// Batched plumes: many independent instances of uneven N in one persistent
// launch, graph-captured leapfrog steps and an autotuned tile shape.
//
// The reference runs one plume per process (DirectForceCalculation::
// updateAcceleration, direct_force_method.cpp:347-428, inside
// LeapfrogDKD::runSimulation, integration_schemes.cpp:856-891); a batch here
// holds the concatenated SoA state of all instances with offsets, and ions of
// one instance interact only with each other.
//
// Targets are taken in Morton order within their instance (so that a tile's
// ions are close to its first ion, the origin of the FP32 pair coordinates);
// the order is formed once per call.  Work list: per instance, tiles of
// WG x TPL targets (WG lanes, TPL targets per lane); items are sorted by their
// source count, largest first, and a
// persistent grid pulls them from an atomic counter (reset by a fill kernel,
// so that the step can be captured in a graph).  Each item streams the
// instance's sources through LDS; pairs are FP32 relative to the tile's first
// target with compensated (two-sum) per-target accumulation (SHP_PREC_FP32C),
// or FP64 (SHP_PREC_FP64).  The autotuner times the candidate (WG, TPL)
// shapes on the caller's state and keeps the fastest.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "shp_internal.h"

struct shp_batch_plan_s {
    shp_handle h = nullptr;
    int ninst = 0, n = 0, prec = SHP_PREC_FP32C;
    std::vector<int> off;
    int cfg = 0;                 // chosen candidate
    int* d_off = nullptr;
    int4* d_items = nullptr;
    int nitems = 0, items_cfg = -1;
    int* d_ctr = nullptr;
    double* d_t = nullptr;
    int* d_eff = nullptr;
    int* d_inst = nullptr;       // instance of each ion
    int* d_order = nullptr;      // ions in (instance, Morton) order
    uint64_t* d_keys = nullptr;  // sort keys (in, out)
    int* d_vals = nullptr;       // sort values (in)
    void* d_cub = nullptr;
    size_t cub_bytes = 0;
    double* d_box = nullptr;     // bounding-box partials and result
    double tuned_ms[8] = {};
};

namespace shp {
namespace {

struct BatchCfg {
    int wg, tpl;
};
constexpr BatchCfg kCfgs[] = {{256, 1}, {256, 2}, {128, 2}, {512, 1}};
constexpr int kNumCfgs = 4;

inline int bdiv2(long a, long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float p) {
    const float s = hi + p;
    const float bp = s - hi;
    lo += (hi - (s - bp)) + (p - bp);
    hi = s;
}

__global__ void k_batch_fill(int* __restrict__ p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

constexpr int kBoxParts = 256;

__global__ __launch_bounds__(256) void k_batch_box(int n, const double* __restrict__ X,
                                                   const double* __restrict__ Y,
                                                   const double* __restrict__ Z,
                                                   double* __restrict__ part) {
    __shared__ double red[6][256];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v[3] = {X[i], Y[i], Z[i]};
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], v[c]);
            mx[c] = fmax(mx[c], v[c]);
        }
    }
    for (int c = 0; c < 3; c++) {
        red[c][threadIdx.x] = mn[c];
        red[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + s]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__device__ __forceinline__ uint32_t bspread10(uint32_t v) {
    v &= 0x3ff;
    v = (v | (v << 16)) & 0x030000ff;
    v = (v | (v << 8)) & 0x0300f00f;
    v = (v | (v << 4)) & 0x030c30c3;
    v = (v | (v << 2)) & 0x09249249;
    return v;
}

// key = instance << 32 | 30-bit Morton code in the batch's bounding box
__global__ void k_batch_keys(int n, const double* __restrict__ X, const double* __restrict__ Y,
                             const double* __restrict__ Z, const int* __restrict__ inst,
                             const double* __restrict__ part, int nparts, uint64_t* __restrict__ keys,
                             int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int b = 0; b < nparts; b++)
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], part[b * 6 + c]);
            mx[c] = fmax(mx[c], part[b * 6 + 3 + c]);
        }
    const double v[3] = {X[i], Y[i], Z[i]};
    uint32_t q[3];
    for (int c = 0; c < 3; c++) {
        const double ext = mx[c] - mn[c];
        const double f = ext > 0.0 ? (v[c] - mn[c]) / ext : 0.0;
        q[c] = (uint32_t)fmin(fmax(f * 1024.0, 0.0), 1023.0);
    }
    keys[i] = ((uint64_t)inst[i] << 32) | (bspread10(q[0]) | (bspread10(q[1]) << 1) | (bspread10(q[2]) << 2));
    vals[i] = i;
}

__global__ void k_batch_inst(int ninst, const int* __restrict__ off, int* __restrict__ inst) {
    const int k = blockIdx.x;
    for (int i = off[k] + threadIdx.x; i < off[k + 1]; i += blockDim.x) inst[i] = k;
}

template <int WG, int TPL, int PREC>
__global__ __launch_bounds__(WG) void k_batch_direct(
    const int4* __restrict__ items, int nitems, int* __restrict__ ctr, const int* __restrict__ off,
    const int* __restrict__ order,
    const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ Z,
    const double* __restrict__ Q, const int* __restrict__ st, const int* __restrict__ sp, double K,
    double eps2, double* __restrict__ AX, double* __restrict__ AY, double* __restrict__ AZ) {
    using Tile = typename std::conditional<PREC == SHP_PREC_FP64, double4, float4>::type;
    __shared__ Tile tile[WG];
    __shared__ int item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item = atomicAdd(ctr, 1);
        __syncthreads();
        const int it = item;
        if (it >= nitems) break;
        const int4 w = items[it];
        const int s0 = off[w.x], s1 = off[w.x + 1];
        const int t0 = w.y, tn = w.z;  // positions in the (instance, Morton) order
        const int io = order[t0];
        const double ox = X[io], oy = Y[io], oz = Z[io];
        int ti[TPL];
        bool tv[TPL];
        double px[TPL], py[TPL], pz[TPL];
        float fhx[TPL], fhy[TPL], fhz[TPL], flx[TPL], fly[TPL], flz[TPL];
        double dhx[TPL], dhy[TPL], dhz[TPL];
#pragma unroll
        for (int k = 0; k < TPL; k++) {
            const int pos = t0 + threadIdx.x + k * WG;
            ti[k] = pos < t0 + tn ? order[pos] : -1;
            tv[k] = ti[k] >= 0 && st[ti[k]] == 1;
            const int i = ti[k] >= 0 ? ti[k] : io;
            px[k] = X[i] - ox;
            py[k] = Y[i] - oy;
            pz[k] = Z[i] - oz;
            fhx[k] = fhy[k] = fhz[k] = flx[k] = fly[k] = flz[k] = 0.0f;
            dhx[k] = dhy[k] = dhz[k] = 0.0;
        }
        for (int j0 = s0; j0 < s1; j0 += WG) {
            __syncthreads();
            const int j = j0 + threadIdx.x;
            const bool src = j < s1 && (st[j] == 1 || st[j] == 2);
            if constexpr (PREC == SHP_PREC_FP64)
                tile[threadIdx.x] = src ? make_double4(X[j] - ox, Y[j] - oy, Z[j] - oz, Q[j])
                                        : make_double4(0.0, 0.0, 0.0, 0.0);
            else
                tile[threadIdx.x] = src ? make_float4((float)(X[j] - ox), (float)(Y[j] - oy),
                                                      (float)(Z[j] - oz), (float)Q[j])
                                        : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            __syncthreads();
            const int cnt = min(WG, s1 - j0);
            if constexpr (PREC == SHP_PREC_FP64) {
#pragma unroll
                for (int k = 0; k < TPL; k++) {
                    double ax = 0, ay = 0, az = 0;
                    for (int m = 0; m < cnt; m++) {
                        const Tile p = tile[m];
                        const double dx = px[k] - (double)p.x, dy = py[k] - (double)p.y, dz = pz[k] - (double)p.z;
                        const double r2 = dx * dx + dy * dy + dz * dz + eps2;
                        const double inv = 1.0 / sqrt(r2);
                        const double ww = (double)p.w * inv * inv * inv;
                        ax += dx * ww;
                        ay += dy * ww;
                        az += dz * ww;
                    }
                    dhx[k] += ax;
                    dhy[k] += ay;
                    dhz[k] += az;
                }
            } else {
                // in a VGPR: an FP32 FMA that reads an SGPR issues at half the
                // rate of one whose operands are all VGPRs
                float e2 = (float)eps2;
                asm volatile("" : "+v"(e2));
                float qx[TPL], qy[TPL], qz[TPL], sx[TPL], sy[TPL], sz[TPL];
#pragma unroll
                for (int k = 0; k < TPL; k++) {
                    qx[k] = (float)px[k];
                    qy[k] = (float)py[k];
                    qz[k] = (float)pz[k];
                    sx[k] = sy[k] = sz[k] = 0.0f;
                }
                for (int m = 0; m < cnt; m++) {
                    const Tile p = tile[m];
#pragma unroll
                    for (int k = 0; k < TPL; k++) {
                        const float dx = qx[k] - (float)p.x, dy = qy[k] - (float)p.y, dz = qz[k] - (float)p.z;
                        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, e2)));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float ww = ((float)p.w * inv) * (inv * inv);
                        sx[k] = fmaf(dx, ww, sx[k]);
                        sy[k] = fmaf(dy, ww, sy[k]);
                        sz[k] = fmaf(dz, ww, sz[k]);
                    }
                }
#pragma unroll
                for (int k = 0; k < TPL; k++) {
                    two_sum_acc(fhx[k], flx[k], sx[k]);
                    two_sum_acc(fhy[k], fly[k], sy[k]);
                    two_sum_acc(fhz[k], flz[k], sz[k]);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < TPL; k++) {
            if (ti[k] < 0) continue;
            const int i = ti[k];
            double ax = 0.0, ay = 0.0, az = 0.0;
            if (tv[k]) {
                const double kc = K * Q[i] / species_mass(sp[i]);
                if (PREC == SHP_PREC_FP64) {
                    ax = kc * dhx[k];
                    ay = kc * dhy[k];
                    az = kc * dhz[k];
                } else {
                    ax = kc * ((double)fhx[k] + (double)flx[k]);
                    ay = kc * ((double)fhy[k] + (double)fly[k]);
                    az = kc * ((double)fhz[k] + (double)flz[k]);
                }
            }
            AX[i] = ax;
            AY[i] = ay;
            AZ[i] = az;
        }
    }
}

// ---------------------------------------------------------------- adjoint ----
// Batched FP64 force VJP (formulas of the single-plume VJP: for an
// acceleration cotangent b, u_j = k_j b_j for targets,
//   xbar_k = sum_j H_kj (q_j u_k - q_k u_j),
//   qbar_k = -sum_j u_j . G_kj + (K/m_k) b_k . sum_j q_j G_kj),
// with j running over the instance of k only; same work list and order.

__global__ void k_batch_vjp_prep(int n, const double* __restrict__ Q, const int* __restrict__ st,
                                 const int* __restrict__ sp, const double* __restrict__ B, double K,
                                 double4* __restrict__ U) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double u0 = 0, u1 = 0, u2 = 0;
    if (st[j] == 1) {
        const double kj = K * Q[j] / species_mass(sp[j]);
        u0 = kj * B[j];
        u1 = kj * B[n + j];
        u2 = kj * B[2 * (size_t)n + j];
    }
    U[j] = make_double4(u0, u1, u2, (st[j] == 1 || st[j] == 2) ? Q[j] : 0.0);
}

template <int WG>
__global__ __launch_bounds__(WG) void k_batch_vjp(
    const int4* __restrict__ items, int nitems, int* __restrict__ ctr, const int* __restrict__ off,
    const int* __restrict__ order, int n, const double* __restrict__ X, const double* __restrict__ Y,
    const double* __restrict__ Z, const int* __restrict__ st, const int* __restrict__ sp,
    const double* __restrict__ B, const double4* __restrict__ U, double K, double eps2,
    double* __restrict__ XB, double* __restrict__ QB) {
    __shared__ double4 tp[WG];
    __shared__ double4 tu[WG];
    __shared__ int item;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) item = atomicAdd(ctr, 1);
        __syncthreads();
        const int it = item;
        if (it >= nitems) break;
        const int4 w = items[it];
        const int s0 = off[w.x], s1 = off[w.x + 1];
        const int pos = w.y + threadIdx.x;
        const int k = pos < w.y + w.z ? order[pos] : -1;
        const bool act = k >= 0 && (st[k] == 1 || st[k] == 2);
        const double xk = act ? X[k] : 0.0, yk = act ? Y[k] : 0.0, zk = act ? Z[k] : 0.0;
        const double4 uk = act ? U[k] : make_double4(0, 0, 0, 0);
        double a0 = 0, a1 = 0, a2 = 0, b0 = 0, b1 = 0, b2 = 0, gu = 0, c0 = 0, c1 = 0, c2 = 0;
        for (int j0 = s0; j0 < s1; j0 += WG) {
            __syncthreads();
            const int j = j0 + threadIdx.x;
            const bool src = j < s1 && (st[j] == 1 || st[j] == 2);
            tp[threadIdx.x] = src ? make_double4(X[j], Y[j], Z[j], 0.0) : make_double4(0, 0, 0, 0);
            tu[threadIdx.x] = src ? U[j] : make_double4(0, 0, 0, 0);
            __syncthreads();
            const int cnt = min(WG, s1 - j0);
            for (int t = 0; t < cnt; t++) {
                if (j0 + t == k) continue;
                const double4 p = tp[t];
                const double4 u = tu[t];
                const double rx = xk - p.x, ry = yk - p.y, rz = zk - p.z;
                const double s = rx * rx + ry * ry + rz * rz + eps2;
                const double is = 1.0 / s;
                const double s32 = is * sqrt(is);
                const double s52 = s32 * is;
                const double rk = rx * uk.x + ry * uk.y + rz * uk.z;
                a0 += u.w * (uk.x * s32 - 3.0 * rx * rk * s52);
                a1 += u.w * (uk.y * s32 - 3.0 * ry * rk * s52);
                a2 += u.w * (uk.z * s32 - 3.0 * rz * rk * s52);
                const double ru = rx * u.x + ry * u.y + rz * u.z;
                b0 += u.x * s32 - 3.0 * rx * ru * s52;
                b1 += u.y * s32 - 3.0 * ry * ru * s52;
                b2 += u.z * s32 - 3.0 * rz * ru * s52;
                gu += ru * s32;
                c0 += u.w * rx * s32;
                c1 += u.w * ry * s32;
                c2 += u.w * rz * s32;
            }
        }
        if (k < 0) continue;
        double rx = 0, ry = 0, rz = 0, rq = 0;
        if (act) {
            rx = a0 - uk.w * b0;
            ry = a1 - uk.w * b1;
            rz = a2 - uk.w * b2;
            rq = -gu;
            if (st[k] == 1) {
                const double km = K / species_mass(sp[k]);
                rq += km * (B[k] * c0 + B[n + k] * c1 + B[2 * (size_t)n + k] * c2);
            }
        }
        XB[k] = rx;
        XB[n + k] = ry;
        XB[2 * (size_t)n + k] = rz;
        QB[k] += rq;
    }
}

// Step VJP pieces (elementwise; same semantics as the single-plume step VJP).
__global__ void k_batch_svjp_pre(int n, const double* __restrict__ X, const double* __restrict__ Y,
                                 const double* __restrict__ Z, const double* __restrict__ VX,
                                 const double* __restrict__ VY, const double* __restrict__ VZ,
                                 const int* __restrict__ st, double dt, const double* __restrict__ XB1,
                                 const double* __restrict__ VB1, double* __restrict__ XH,
                                 double* __restrict__ W, double* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = st[i] == 1 || st[i] == 2;
    const double h = 0.5 * dt;
    const double vv[3] = {VX[i], VY[i], VZ[i]};
    const double xx[3] = {X[i], Y[i], Z[i]};
    for (int c = 0; c < 3; c++) {
        const size_t o = (size_t)c * n + i;
        XH[o] = mv ? xx[c] + vv[c] * h : xx[c];
        const double w = mv ? VB1[o] + XB1[o] * h : VB1[o];
        W[o] = w;
        B[o] = st[i] == 1 ? dt * w : 0.0;
    }
}

__global__ void k_batch_svjp_post(int n, const int* __restrict__ st, double dt,
                                  const double* __restrict__ XB1, const double* __restrict__ W,
                                  double* __restrict__ XB0, double* __restrict__ VB0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = st[i] == 1 || st[i] == 2;
    const double h = mv ? 0.5 * dt : 0.0;
    for (int c = 0; c < 3; c++) {
        const size_t o = (size_t)c * n + i;
        const double x0 = XB1[o] + XB0[o];
        XB0[o] = x0;
        VB0[o] = W[o] + x0 * h;
    }
}

template <int WG, int TPL, int PREC>
void launch_cfg(shp_batch_plan_s* b, const shp_particles* p, double* ax, double* ay, double* az) {
    const int grid = b->h->num_cus * std::max(1, 1024 / WG);
    hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(64), 0, b->h->stream, b->d_ctr, 0);
    hipLaunchKernelGGL((k_batch_direct<WG, TPL, PREC>), dim3(grid), dim3(WG), 0, b->h->stream, b->d_items,
                       b->nitems, b->d_ctr, b->d_off, b->d_order, p->x, p->y, p->z, p->q, p->status,
                       p->species,
                       k_const_host(), kSoftening, ax, ay, az);
}

template <int PREC>
void launch_prec(shp_batch_plan_s* b, int cfg, const shp_particles* p, double* ax, double* ay, double* az) {
    switch (cfg) {
    case 0: launch_cfg<256, 1, PREC>(b, p, ax, ay, az); break;
    case 1: launch_cfg<256, 2, PREC>(b, p, ax, ay, az); break;
    case 2: launch_cfg<128, 2, PREC>(b, p, ax, ay, az); break;
    default: launch_cfg<512, 1, PREC>(b, p, ax, ay, az); break;
    }
}

int upload_items(shp_batch_plan_s* b, int cfg) {
    if (b->items_cfg == cfg) return SHP_OK;
    const int per = kCfgs[cfg].wg * kCfgs[cfg].tpl;
    std::vector<int4> it;
    for (int k = 0; k < b->ninst; k++)
        for (int t = b->off[k]; t < b->off[k + 1]; t += per)
            it.push_back(make_int4(k, t, std::min(per, b->off[k + 1] - t), b->off[k + 1] - b->off[k]));
    std::stable_sort(it.begin(), it.end(), [](const int4& a, const int4& c) { return a.w > c.w; });
    if (b->d_items) (void)hipFree(b->d_items);
    b->d_items = nullptr;
    SHP_HIP_CHECK(hipMalloc((void**)&b->d_items, sizeof(int4) * std::max<size_t>(1, it.size())));
    if (!it.empty())
        SHP_HIP_CHECK(hipMemcpy(b->d_items, it.data(), sizeof(int4) * it.size(), hipMemcpyHostToDevice));
    b->nitems = (int)it.size();
    b->items_cfg = cfg;
    return SHP_OK;
}

// Target order: Morton within each instance, from the current positions.
int batch_order(shp_batch_plan_s* b, const double* X, const double* Y, const double* Z) {
    const int n = b->n;
    hipStream_t st = b->h->stream;
    hipLaunchKernelGGL(k_batch_box, dim3(kBoxParts), dim3(256), 0, st, n, X, Y, Z, b->d_box);
    hipLaunchKernelGGL(k_batch_keys, dim3(bdiv2(n, 256)), dim3(256), 0, st, n, X, Y, Z, b->d_inst, b->d_box,
                       kBoxParts, b->d_keys, b->d_vals);
    size_t tb = b->cub_bytes;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(b->d_cub, tb, b->d_keys, b->d_keys + n, b->d_vals,
                                                     b->d_order, n, 0, 64, st));
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

int batch_forces(shp_batch_plan_s* b, int cfg, const shp_particles* p, double* ax, double* ay, double* az) {
    int rc = upload_items(b, cfg);
    if (rc) return rc;
    if (b->prec == SHP_PREC_FP64) launch_prec<SHP_PREC_FP64>(b, cfg, p, ax, ay, az);
    else launch_prec<SHP_PREC_FP32C>(b, cfg, p, ax, ay, az);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

}  // namespace
}  // namespace shp

using namespace shp;

extern "C" {

int shp_batch_create(shp_handle h, int ninst, const int* offsets, int precision, shp_batch_plan* out) {
    SHP_REQUIRE(h && offsets && out, "shp_batch_create: null argument");
    SHP_REQUIRE(ninst >= 1, "shp_batch_create: ninst=%d", ninst);
    SHP_REQUIRE(precision == SHP_PREC_FP32C || precision == SHP_PREC_FP64,
                "shp_batch_create: precision %d (FP32C or FP64)", precision);
    SHP_REQUIRE(offsets[0] == 0, "shp_batch_create: offsets[0] must be 0");
    for (int k = 0; k < ninst; k++)
        SHP_REQUIRE(offsets[k + 1] >= offsets[k], "shp_batch_create: offsets not ascending at %d", k);
    SHP_HIP_CHECK(hipSetDevice(h->device));
    auto* b = new shp_batch_plan_s();
    b->h = h;
    b->ninst = ninst;
    b->n = offsets[ninst];
    b->prec = precision;
    b->off.assign(offsets, offsets + ninst + 1);
    hipError_t e = hipMalloc((void**)&b->d_off, sizeof(int) * (ninst + 1));
    if (e == hipSuccess) e = hipMemcpy(b->d_off, offsets, sizeof(int) * (ninst + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_ctr, sizeof(int));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_t, sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_eff, sizeof(int) * std::max(1, b->n));
    const size_t nn = std::max(1, b->n);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_inst, sizeof(int) * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_order, sizeof(int) * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_keys, sizeof(uint64_t) * 2 * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_vals, sizeof(int) * nn);
    if (e == hipSuccess) e = hipMalloc((void**)&b->d_box, sizeof(double) * 6 * kBoxParts);
    if (e == hipSuccess)
        e = hipcub::DeviceRadixSort::SortPairs(nullptr, b->cub_bytes, b->d_keys, b->d_keys + nn, b->d_vals,
                                               b->d_order, (int)nn, 0, 64, h->stream);
    if (e == hipSuccess) e = hipMalloc(&b->d_cub, std::max<size_t>(b->cub_bytes, 16));
    if (e == hipSuccess && b->n > 0) {
        hipLaunchKernelGGL(k_batch_inst, dim3(ninst), dim3(256), 0, h->stream, ninst, b->d_off, b->d_inst);
        e = hipStreamSynchronize(h->stream);
    }
    if (e != hipSuccess) {
        set_error("shp_batch_create: %s", hipGetErrorString(e));
        shp_batch_destroy(b);
        return SHP_ERR_HIP;
    }
    *out = b;
    return SHP_OK;
}

int shp_batch_destroy(shp_batch_plan b) {
    if (!b) return SHP_OK;
    (void)hipDeviceSynchronize();
    void* ptrs[] = {b->d_off, b->d_items, b->d_ctr, b->d_t, b->d_eff, b->d_inst, b->d_order,
                    b->d_keys, b->d_vals, b->d_cub, b->d_box};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
    return SHP_OK;
}

int shp_batch_tune(shp_batch_plan b, const shp_particles* all, double* ms_out, int* chosen) {
    SHP_REQUIRE(b && all, "shp_batch_tune: null argument");
    SHP_REQUIRE(all->n == b->n, "shp_batch_tune: n=%d, plan holds %d ions", all->n, b->n);
    shp_handle h = b->h;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(scr, double, h, WS_BATCH_TUNE, 3 * (size_t)std::max(1, b->n));
    int orc = batch_order(b, all->x, all->y, all->z);
    if (orc) return orc;
    double best = 1e300;
    int arg = 0;
    for (int c = 0; c < kNumCfgs; c++) {
        int rc = batch_forces(b, c, all, scr, scr + b->n, scr + 2 * (size_t)b->n);  // warm-up
        if (rc) return rc;
        SHP_HIP_CHECK(hipEventRecord(h->ev0, h->stream));
        for (int r = 0; r < 3; r++) {
            rc = batch_forces(b, c, all, scr, scr + b->n, scr + 2 * (size_t)b->n);
            if (rc) return rc;
        }
        SHP_HIP_CHECK(hipEventRecord(h->ev1, h->stream));
        SHP_HIP_CHECK(hipEventSynchronize(h->ev1));
        float ms = 0.0f;
        SHP_HIP_CHECK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        b->tuned_ms[c] = ms / 3.0;
        if (ms_out) ms_out[c] = ms / 3.0;
        if (ms < best) {
            best = ms;
            arg = c;
        }
    }
    b->cfg = arg;
    if (chosen) *chosen = arg;
    return SHP_OK;
}

int shp_batch_set_config(shp_batch_plan b, int cfg) {
    SHP_REQUIRE(b && cfg >= 0 && cfg < kNumCfgs, "shp_batch_set_config: config %d of %d", cfg, kNumCfgs);
    b->cfg = cfg;
    return SHP_OK;
}

int shp_batch_forces(shp_batch_plan b, const shp_particles* all, double* ax, double* ay, double* az) {
    SHP_REQUIRE(b && all && ax && ay && az, "shp_batch_forces: null argument");
    SHP_REQUIRE(all->n == b->n, "shp_batch_forces: n=%d, plan holds %d ions", all->n, b->n);
    SHP_HIP_CHECK(hipSetDevice(b->h->device));
    if (b->n == 0) return SHP_OK;
    int rc = batch_order(b, all->x, all->y, all->z);
    if (rc) return rc;
    SHP_HIP_CHECK(hipEventRecord(b->h->ev0, b->h->stream));
    rc = batch_forces(b, b->cfg, all, ax, ay, az);
    SHP_HIP_CHECK(hipEventRecord(b->h->ev1, b->h->stream));
    b->h->timed = true;
    return rc;
}

int shp_batch_leapfrog(shp_batch_plan b, shp_state* s, const shp_step_opts* o, double t0, double dt,
                       int nsteps, int use_graph, double* ax, double* ay, double* az) {
    SHP_REQUIRE(b && s && o && ax && ay && az, "shp_batch_leapfrog: null argument");
    SHP_REQUIRE(s->n == b->n, "shp_batch_leapfrog: n=%d, plan holds %d ions", s->n, b->n);
    SHP_REQUIRE(o->sym.type == SHP_SYM_NONE, "shp_batch_leapfrog: instances have no symmetry images");
    SHP_REQUIRE(nsteps >= 0 && dt >= 0, "shp_batch_leapfrog: steps=%d dt=%g", nsteps, dt);
    shp_handle h = b->h;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    if (b->n == 0 || nsteps == 0) return SHP_OK;
    int rc = upload_items(b, b->cfg);  // outside any capture
    if (rc) return rc;
    rc = lf_begin(h, s, o, t0, dt, b->d_eff);
    if (rc) return rc;
    rc = batch_order(b, s->x, s->y, s->z);  // kept for the steps of this call
    if (rc) return rc;
    SHP_HIP_CHECK(hipMemcpyAsync(b->d_t, &t0, sizeof(double), hipMemcpyHostToDevice, h->stream));
    shp_particles p{s->n, s->x, s->y, s->z, s->q, b->d_eff, s->species};
    auto step = [&](int next) -> int {
        int r = batch_forces(b, b->cfg, &p, ax, ay, az);
        if (r) return r;
        return lf_kick_drift(h, s, o, 0.0, b->d_t, dt, next, ax, ay, az, b->d_eff);
    };
    int done = 0;
    if (use_graph && nsteps > 1) {
        // capture one regular step (forces, kick, drifts, time advance) and
        // replay it; the last step runs uncaptured without the next drift
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        SHP_HIP_CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        const int r = step(1);
        hipError_t e = hipStreamEndCapture(h->stream, &graph);
        if (r) {
            if (graph) (void)hipGraphDestroy(graph);
            return r;
        }
        SHP_HIP_CHECK(e);
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            SHP_HIP_CHECK(e);
        }
        for (; done < nsteps - 1; done++) {
            e = hipGraphLaunch(exec, h->stream);
            if (e != hipSuccess) break;
        }
        (void)hipStreamSynchronize(h->stream);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        SHP_HIP_CHECK(e);
    }
    for (; done < nsteps; done++) {
        rc = step(done + 1 < nsteps ? 1 : 0);
        if (rc) return rc;
    }
    return SHP_OK;
}

// Cotangent arrays are component-major: xb1[c * n + i] (c = x, y, z), n
// the batch's ion count.
int shp_batch_step_vjp(shp_batch_plan b, const shp_state* s, double dt, const double* xb1,
                       const double* vb1, double* xb0, double* vb0, double* qb) {
    SHP_REQUIRE(b && s && xb1 && vb1 && xb0 && vb0 && qb, "shp_batch_step_vjp: null argument");
    SHP_REQUIRE(s->n == b->n, "shp_batch_step_vjp: n=%d, plan holds %d ions", s->n, b->n);
    shp_handle h = b->h;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    const int n = b->n;
    if (n == 0) return SHP_OK;
    constexpr int WG = 256;
    // work list of WG targets per item (the FP64 configuration 0 shape)
    int rc = upload_items(b, 0);
    if (rc) return rc;
    rc = batch_order(b, s->x, s->y, s->z);
    if (rc) return rc;
    SHP_WS(tmp, double, h, WS_ADJ_TMP, 9 * (size_t)n);
    SHP_WS(U, double4, h, WS_ADJ_U, n);
    double *XH = tmp, *W = tmp + 3 * (size_t)n, *B = tmp + 6 * (size_t)n;
    const dim3 g(bdiv2(n, 256)), blk(256);
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(k_batch_svjp_pre, g, blk, 0, st, n, s->x, s->y, s->z, s->vx, s->vy, s->vz,
                       s->status, dt, xb1, vb1, XH, W, B);
    hipLaunchKernelGGL(k_batch_vjp_prep, g, blk, 0, st, n, s->q, s->status, s->species, B, k_const_host(), U);
    hipLaunchKernelGGL(k_batch_fill, dim3(1), dim3(64), 0, st, b->d_ctr, 0);
    hipLaunchKernelGGL(k_batch_vjp<WG>, dim3(h->num_cus * 4), dim3(WG), 0, st, b->d_items, b->nitems,
                       b->d_ctr, b->d_off, b->d_order, n, XH, XH + n, XH + 2 * (size_t)n, s->status,
                       s->species, B, U, k_const_host(), kSoftening, xb0, qb);
    hipLaunchKernelGGL(k_batch_svjp_post, g, blk, 0, st, n, s->status, dt, xb1, W, xb0, vb0);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

}  // extern "C"
