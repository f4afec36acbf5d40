// Device kernels of the direct P2P sums (included by direct.hip only).
#pragma once

#include "shp_internal.h"

namespace shp {

constexpr int kThreads = 256;
constexpr int kTile = 256;
constexpr int kF32TargetsPerLane = 2;

// ---------------------------------------------------------------- staging ----

__global__ void k_flag_sources_f64(const int* __restrict__ status, int j0, int n, char* flags) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        int s = status[j0 + t];
        flags[t] = (s == 1 || s == 2) ? 1 : 0;
    }
}

__global__ void k_flag_sources_ref(const float4* __restrict__ acc4, int j0, int n, char* flags) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) {
        float s = acc4[j0 + t].w;
        flags[t] = (s == 1.0f || s == 2.0f) ? 1 : 0;
    }
}

__global__ void k_gather_sources_f64(const int* __restrict__ idx, const int* __restrict__ count,
                                     const double* __restrict__ X, const double* __restrict__ Y,
                                     const double* __restrict__ Z, const double* __restrict__ Q,
                                     int cap, double4* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cap && t < *count) {
        int j = idx[t];
        out[t] = make_double4(X[j], Y[j], Z[j], Q[j]);
    }
}

__global__ void k_gather_sources_ref(const int* __restrict__ idx, const int* __restrict__ count,
                                     const float4* __restrict__ pos4, int cap,
                                     float4* __restrict__ out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < cap && t < *count) out[t] = pos4[idx[t]];
}

// ------------------------------------------------------ Morton target order ----

__global__ void k_bbox_partial(const double* __restrict__ X, const double* __restrict__ Y,
                               const double* __restrict__ Z, const int* __restrict__ status,
                               int i0, int n, double* __restrict__ part) {
    __shared__ double red[6][kThreads];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        int i = i0 + t;
        if (status[i] != 1) continue;
        double v[3] = {X[i], Y[i], Z[i]};
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
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            for (int c = 0; c < 3; c++) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + s]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_bbox_final(const double* __restrict__ part, int nparts, double* __restrict__ box) {
    if (threadIdx.x < 6) {
        double v = threadIdx.x < 3 ? 1e300 : -1e300;
        for (int b = 0; b < nparts; b++) {
            double p = part[b * 6 + threadIdx.x];
            v = threadIdx.x < 3 ? fmin(v, p) : fmax(v, p);
        }
        box[threadIdx.x] = v;
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void k_morton30(const double* __restrict__ X, const double* __restrict__ Y,
                           const double* __restrict__ Z, const int* __restrict__ status, int i0,
                           int n, const double* __restrict__ box, uint32_t* __restrict__ keys,
                           int* __restrict__ vals) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int i = i0 + t;
    uint32_t key = 0xffffffffu;
    if (status[i] == 1) {
        double ext = fmax(fmax(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
        double sc = ext > 0 ? 1023.999 / ext : 0.0;
        uint32_t ix = (uint32_t)fmin(fmax((X[i] - box[0]) * sc, 0.0), 1023.0);
        uint32_t iy = (uint32_t)fmin(fmax((Y[i] - box[1]) * sc, 0.0), 1023.0);
        uint32_t iz = (uint32_t)fmin(fmax((Z[i] - box[2]) * sc, 0.0), 1023.0);
        key = spread10(ix) | (spread10(iy) << 1) | (spread10(iz) << 2);
    }
    keys[t] = key;
    vals[t] = i;
}

__global__ void k_iota(int* __restrict__ out, int base, int n) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = base + t;
}

__global__ void k_fill(double* __restrict__ p, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

__global__ void k_aos_to_f64(const float4* __restrict__ pos4, const float4* __restrict__ vel4,
                             const float4* __restrict__ acc4, int n, double* X, double* Y,
                             double* Z, double* Q, int* status, int* species) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pos4[i];
    X[i] = p.x;
    Y[i] = p.y;
    Z[i] = p.z;
    Q[i] = p.w;
    status[i] = (int)acc4[i].w;
    species[i] = (int)vel4[i].w;
}

// -------------------------------------------------------------- P2P kernels ----

__device__ __forceinline__ void split_range(int nsrc, int nsplit, int split, int& jb, int& je) {
    int chunk = (nsrc + nsplit - 1) / nsplit;
    chunk = (chunk + kTile - 1) / kTile * kTile;
    jb = split * chunk;
    je = min(nsrc, jb + chunk);
}

// FP64 state: double positions, double pair arithmetic, one target per lane.
template <int SYM>
__global__ __launch_bounds__(kThreads) void k_direct_f64(
    const int* __restrict__ tgt, int nslots, const double* __restrict__ X,
    const double* __restrict__ Y, const double* __restrict__ Z, const double4* __restrict__ src,
    const int* __restrict__ nsrc_p, const double4* __restrict__ img, int nimg, int nsplit,
    double eps, double* __restrict__ part, int cap) {
    __shared__ double4 tile[kTile];
    const int slot = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = slot < nslots;
    const int i = tgt[valid ? slot : 0];
    const double xi = X[i], yi = Y[i], zi = Z[i];
    int jb, je;
    split_range(*nsrc_p, nsplit, blockIdx.y, jb, je);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int m = 0; m < nimg; m++) {
        const double4 tr = img[m];
        double txi = xi, tyi = yi, tzi = zi;
        if (SYM == SHP_SYM_TRANSLATIONAL) {
            txi -= tr.x;
            tyi -= tr.y;
            tzi -= tr.z;
        }
        for (int base = jb; base < je; base += kTile) {
            __syncthreads();
            const int j = base + threadIdx.x;
            double4 s = j < je ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
            if (SYM == SHP_SYM_ROTATIONAL) {
                double rx = tr.x * s.x - tr.y * s.y;
                double ry = tr.y * s.x + tr.x * s.y;
                s.x = rx;
                s.y = ry;
            }
            tile[threadIdx.x] = s;
            __syncthreads();
            const int cnt = min(kTile, je - base);
#pragma unroll 4
            for (int jj = 0; jj < cnt; jj++) {
                const double4 sj = tile[jj];
                const double dx = txi - sj.x, dy = tyi - sj.y, dz = tzi - sj.z;
                const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
                const double inv = rsqrt_f64(r2);
                const double w = sj.w * inv * inv * inv;
                ax = fma(dx, w, ax);
                ay = fma(dy, w, ay);
                az = fma(dz, w, az);
            }
        }
    }
    if (valid) {
        const size_t o = (size_t)blockIdx.y * 3 * cap + slot;
        part[o] = ax;
        part[o + cap] = ay;
        part[o + 2 * (size_t)cap] = az;
    }
}

// Reference contract on float AoS positions: float displacements (and float
// source rotation / target shift), double distance, rsqrt and sums.
template <int SYM>
__global__ __launch_bounds__(kThreads) void k_direct_ref(
    const int* __restrict__ tgt, int nslots, const float4* __restrict__ pos4,
    const float4* __restrict__ src, const int* __restrict__ nsrc_p, const double4* __restrict__ img,
    int nimg, int nsplit, double eps, double* __restrict__ part, int cap) {
#pragma clang fp contract(off)
    __shared__ float4 tile[kTile];
    const int slot = blockIdx.x * kThreads + threadIdx.x;
    const bool valid = slot < nslots;
    const float4 pi = pos4[tgt[valid ? slot : 0]];
    int jb, je;
    split_range(*nsrc_p, nsplit, blockIdx.y, jb, je);
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int m = 0; m < nimg; m++) {
        const double4 tr = img[m];
        const float c = (float)tr.x, s = (float)tr.y;
        float txi = pi.x, tyi = pi.y, tzi = pi.z;
        if (SYM == SHP_SYM_TRANSLATIONAL) {
            txi = pi.x - (float)tr.x;
            tyi = pi.y - (float)tr.y;
            tzi = pi.z - (float)tr.z;
        }
        for (int base = jb; base < je; base += kTile) {
            __syncthreads();
            const int j = base + threadIdx.x;
            float4 sj = j < je ? src[j] : make_float4(0.f, 0.f, 0.f, 0.f);
            if (SYM == SHP_SYM_ROTATIONAL) {
                const float cx = c * sj.x, sy = s * sj.y, sx = s * sj.x, cy = c * sj.y;
                sj.x = cx - sy;
                sj.y = sx + cy;
            }
            tile[threadIdx.x] = sj;
            __syncthreads();
            const int cnt = min(kTile, je - base);
#pragma unroll 4
            for (int jj = 0; jj < cnt; jj++) {
                const float4 q = tile[jj];
                const double dx = (double)(txi - q.x);
                const double dy = (double)(tyi - q.y);
                const double dz = (double)(tzi - q.z);
                const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps)));
                const double inv = rsqrt_f64(r2);
                const double w = (double)q.w * inv * inv * inv;
                ax = fma(dx, w, ax);
                ay = fma(dy, w, ay);
                az = fma(dz, w, az);
            }
        }
    }
    if (valid) {
        const size_t o = (size_t)blockIdx.y * 3 * cap + slot;
        part[o] = ax;
        part[o + cap] = ay;
        part[o + 2 * (size_t)cap] = az;
    }
}

// Knuth two-sum of a partial into a (hi, lo) pair.
__device__ __forceinline__ void two_sum_acc(float& hi, float& lo, float p) {
    const float s = hi + p;
    const float bp = s - hi;
    const float err = (hi - (s - bp)) + (p - bp);
    hi = s;
    lo += err;
}

// FP32 pair arithmetic with compensated per-target accumulation; two targets
// per lane, Morton-ordered target slots, per-workgroup origin.
template <int SYM>
__global__ __launch_bounds__(kThreads) void k_direct_f32c(
    const int* __restrict__ tgt, int nslots, const double* __restrict__ X,
    const double* __restrict__ Y, const double* __restrict__ Z, const double4* __restrict__ src,
    const int* __restrict__ nsrc_p, const double4* __restrict__ img, int nimg, int nsplit,
    float eps, double* __restrict__ part, int cap) {
    constexpr int T = kF32TargetsPerLane;
    __shared__ float4 tile[kTile];
    const int slot0 = blockIdx.x * (kThreads * T);
    const int org = tgt[slot0];
    const double ox = X[org], oy = Y[org], oz = Z[org];
    float tx[T], ty[T], tz[T];
    float hx[T], hy[T], hz[T], lx[T], ly[T], lz[T];
#pragma unroll
    for (int k = 0; k < T; k++) {
        const int slot = slot0 + threadIdx.x + k * kThreads;
        const int i = tgt[slot < nslots ? slot : slot0];
        tx[k] = (float)(X[i] - ox);
        ty[k] = (float)(Y[i] - oy);
        tz[k] = (float)(Z[i] - oz);
        hx[k] = hy[k] = hz[k] = lx[k] = ly[k] = lz[k] = 0.0f;
    }
    int jb, je;
    split_range(*nsrc_p, nsplit, blockIdx.y, jb, je);
    // eps in a VGPR: an FP32 FMA that reads an SGPR issues at half the rate of
    // one whose operands are all VGPRs.
    float epsv = eps;
    asm volatile("" : "+v"(epsv));
    for (int m = 0; m < nimg; m++) {
        const double4 tr = img[m];
        for (int base = jb; base < je; base += kTile) {
            __syncthreads();
            const int j = base + threadIdx.x;
            float4 st = make_float4(0.f, 0.f, 0.f, 0.f);
            if (j < je) {
                double4 s = src[j];
                if (SYM == SHP_SYM_ROTATIONAL) {
                    const double rx = tr.x * s.x - tr.y * s.y;
                    const double ry = tr.y * s.x + tr.x * s.y;
                    s.x = rx;
                    s.y = ry;
                } else if (SYM == SHP_SYM_TRANSLATIONAL) {
                    s.x += tr.x;
                    s.y += tr.y;
                    s.z += tr.z;
                }
                st = make_float4((float)(s.x - ox), (float)(s.y - oy), (float)(s.z - oz), (float)s.w);
            }
            tile[threadIdx.x] = st;
            __syncthreads();
            const int cnt = min(kTile, je - base);
            float px[T], py[T], pz[T];
#pragma unroll
            for (int k = 0; k < T; k++) px[k] = py[k] = pz[k] = 0.0f;
#pragma unroll 8
            for (int jj = 0; jj < cnt; jj++) {
                const float4 q = tile[jj];
#pragma unroll
                for (int k = 0; k < T; k++) {
                    const float dx = tx[k] - q.x, dy = ty[k] - q.y, dz = tz[k] - q.z;
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, epsv)));
                    const float inv = __builtin_amdgcn_rsqf(r2);
                    const float w = (q.w * inv) * (inv * inv);
                    px[k] = fmaf(dx, w, px[k]);
                    py[k] = fmaf(dy, w, py[k]);
                    pz[k] = fmaf(dz, w, pz[k]);
                }
            }
#pragma unroll
            for (int k = 0; k < T; k++) {
                two_sum_acc(hx[k], lx[k], px[k]);
                two_sum_acc(hy[k], ly[k], py[k]);
                two_sum_acc(hz[k], lz[k], pz[k]);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < T; k++) {
        const int slot = slot0 + threadIdx.x + k * kThreads;
        if (slot < nslots) {
            const size_t o = (size_t)blockIdx.y * 3 * cap + slot;
            part[o] = (double)hx[k] + (double)lx[k];
            part[o + cap] = (double)hy[k] + (double)ly[k];
            part[o + 2 * (size_t)cap] = (double)hz[k] + (double)lz[k];
        }
    }
}

// Combine split partials in split order, scale by K q_i / m_i, scatter.
// mode 0: SoA device output indexed by particle; mode 1: AoS rows (i - i0).
template <bool REF>
__global__ void k_reduce_scatter(const int* __restrict__ tgt, int nslots, int nsplit,
                                 const double* __restrict__ part, int cap,
                                 const int* __restrict__ status, const double* __restrict__ Q,
                                 const int* __restrict__ species, const float4* __restrict__ pos4,
                                 const float4* __restrict__ vel4, const float4* __restrict__ acc4,
                                 double K, int mode, int i0, double* __restrict__ ax,
                                 double* __restrict__ ay, double* __restrict__ az,
                                 double* __restrict__ out3) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= nslots) return;
    const int i = tgt[slot];
    double kc;
    if (REF) {
        if (acc4[i].w != 1.0f) return;
        kc = K * (double)pos4[i].w / species_mass((int)vel4[i].w);
    } else {
        if (status[i] != 1) return;
        kc = K * Q[i] / species_mass(species[i]);
    }
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int s = 0; s < nsplit; s++) {
        const size_t o = (size_t)s * 3 * cap + slot;
        sx += part[o];
        sy += part[o + cap];
        sz += part[o + 2 * (size_t)cap];
    }
    if (mode == 0) {
        ax[i] = sx * kc;
        ay[i] = sy * kc;
        az[i] = sz * kc;
    } else {
        double* r = out3 + 3 * (size_t)(i - i0);
        r[0] = sx * kc;
        r[1] = sy * kc;
        r[2] = sz * kc;
    }
}

}  // namespace shp
