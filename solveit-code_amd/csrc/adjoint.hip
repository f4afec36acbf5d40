// Derivatives of the FP64 direct force and of the DKD leapfrog step.
//
// The reference has no derivatives; these differentiate the engine's FP64
// direct sum (shp_direct_dev with SHP_PREC_FP64, direct_force_method.cpp:153-190
// in FP64) and its DKD step (shp_leapfrog_run; integration_schemes.cpp:856-891)
// for ions that are all injected, without background field or boundary
// conditions.
//
//   a_i = k_i sum_j q_j G(r_i - r_j),   k_i = K q_i / m_i,
//   G(r) = r s^-3/2,  H(r) = dG/dr = I s^-3/2 - 3 r r^T s^-5/2,  s = |r|^2 + eps^2.
//
// Force JVP (tangents dx, dq):
//   da_i = k_i sum_j q_j H_ij (dx_i - dx_j) + k_i sum_j dq_j G_ij + (K dq_i / m_i) sum_j q_j G_ij
// Force VJP (cotangent b of a; u_j = k_j b_j for targets, 0 otherwise):
//   xbar_k = sum_j H_kj (q_j u_k - q_k u_j)
//   qbar_k = -sum_j u_j . G_kj + (K / m_k) b_k . sum_j q_j G_kj
// Both are N-body sums with the rank-2 kernel H, evaluated like the direct
// sum: a workgroup of targets streams LDS tiles of sources.
//
// DKD step and its VJP (cotangents xb1, vb1 of the step's end state):
//   xh = x + v dt/2,  v1 = v + dt a(xh),  x1 = xh + v1 dt/2
//   w = vb1 + xb1 dt/2,  b = dt w,  xhb = xb1 + J_x^T b,  qb += J_q^T b,
//   xb0 = xhb,  vb0 = w + xhb dt/2
// Status 1 ions drift and kick, status 2 ions drift only (sources), others
// are frozen.
#include <algorithm>
#include <cmath>

#include "shp_internal.h"

namespace shp {
namespace {

constexpr int kAdjThreads = 256;

inline int adiv(long a, long b) { return (int)((a + b - 1) / b); }

__device__ __forceinline__ bool is_src(int s) { return s == 1 || s == 2; }

__global__ __launch_bounds__(kAdjThreads) void k_force_jvp(
    int n, const double* __restrict__ X, const double* __restrict__ Y,
    const double* __restrict__ Z, const double* __restrict__ Q, const int* __restrict__ st,
    const int* __restrict__ sp, const double* __restrict__ DX, const double* __restrict__ DY,
    const double* __restrict__ DZ, const double* __restrict__ DQ, double K, double eps2,
    double* __restrict__ oax, double* __restrict__ oay, double* __restrict__ oaz) {
    __shared__ double4 tp[kAdjThreads];  // x, y, z, q (0 for non-sources)
    __shared__ double4 td[kAdjThreads];  // dx, dy, dz, dq (0 for non-sources)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool tgt = i < n && st[i] == 1;
    const double xi = tgt ? X[i] : 0.0, yi = tgt ? Y[i] : 0.0, zi = tgt ? Z[i] : 0.0;
    const double dxi = tgt && DX ? DX[i] : 0.0, dyi = tgt && DY ? DY[i] : 0.0,
                 dzi = tgt && DZ ? DZ[i] : 0.0;
    double h0 = 0, h1 = 0, h2 = 0, g0 = 0, g1 = 0, g2 = 0, c0 = 0, c1 = 0, c2 = 0;
    for (int j0 = 0; j0 < n; j0 += kAdjThreads) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        const bool src = j < n && is_src(st[j]);
        tp[threadIdx.x] = src ? make_double4(X[j], Y[j], Z[j], Q[j]) : make_double4(0, 0, 0, 0);
        td[threadIdx.x] = src ? make_double4(DX ? DX[j] : 0.0, DY ? DY[j] : 0.0, DZ ? DZ[j] : 0.0,
                                             DQ ? DQ[j] : 0.0)
                              : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = min(kAdjThreads, n - j0);
        for (int k = 0; k < cnt; k++) {
            if (j0 + k == i) continue;
            const double4 p = tp[k];
            const double4 dd = td[k];
            const double rx = xi - p.x, ry = yi - p.y, rz = zi - p.z;
            const double s = rx * rx + ry * ry + rz * rz + eps2;
            const double is = 1.0 / s;
            const double s32 = is * sqrt(is);
            const double s52 = s32 * is;
            const double ex = dxi - dd.x, ey = dyi - dd.y, ez = dzi - dd.z;
            const double rd = rx * ex + ry * ey + rz * ez;
            // q_j H (dx_i - dx_j)
            h0 += p.w * (ex * s32 - 3.0 * rx * rd * s52);
            h1 += p.w * (ey * s32 - 3.0 * ry * rd * s52);
            h2 += p.w * (ez * s32 - 3.0 * rz * rd * s52);
            // dq_j G and q_j G
            g0 += dd.w * rx * s32;
            g1 += dd.w * ry * s32;
            g2 += dd.w * rz * s32;
            c0 += p.w * rx * s32;
            c1 += p.w * ry * s32;
            c2 += p.w * rz * s32;
        }
    }
    if (i < n) {
        if (tgt) {
            const double m = species_mass(sp[i]);
            const double ki = K * Q[i] / m;
            const double kd = K * (DQ ? DQ[i] : 0.0) / m;
            oax[i] = ki * (h0 + g0) + kd * c0;
            oay[i] = ki * (h1 + g1) + kd * c1;
            oaz[i] = ki * (h2 + g2) + kd * c2;
        } else {
            oax[i] = oay[i] = oaz[i] = 0.0;
        }
    }
}

// u_j = k_j b_j for targets (0 otherwise), packed with q_j.
__global__ void k_vjp_prep(int n, const double* __restrict__ Q, const int* __restrict__ st,
                           const int* __restrict__ sp, const double* __restrict__ BX,
                           const double* __restrict__ BY, const double* __restrict__ BZ, double K,
                           double4* __restrict__ U) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double u0 = 0, u1 = 0, u2 = 0;
    if (st[j] == 1) {
        const double kj = K * Q[j] / species_mass(sp[j]);
        u0 = kj * BX[j];
        u1 = kj * BY[j];
        u2 = kj * BZ[j];
    }
    U[j] = make_double4(u0, u1, u2, is_src(st[j]) ? Q[j] : 0.0);
}

__global__ __launch_bounds__(kAdjThreads) void k_force_vjp(
    int n, const double* __restrict__ X, const double* __restrict__ Y,
    const double* __restrict__ Z, const int* __restrict__ st, const int* __restrict__ sp,
    const double* __restrict__ BX, const double* __restrict__ BY, const double* __restrict__ BZ,
    const double4* __restrict__ U, double K, double eps2, double* __restrict__ xb,
    double* __restrict__ yb, double* __restrict__ zb, double* __restrict__ qb, int acc_x, int acc_q) {
    __shared__ double4 tp[kAdjThreads];  // x, y, z (sources)
    __shared__ double4 tu[kAdjThreads];  // u, q
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = k < n && is_src(st[k]);
    const double xk = act ? X[k] : 0.0, yk = act ? Y[k] : 0.0, zk = act ? Z[k] : 0.0;
    const double4 uk = act ? U[k] : make_double4(0, 0, 0, 0);
    double a0 = 0, a1 = 0, a2 = 0;      // sum_j q_j H_kj u_k
    double b0 = 0, b1 = 0, b2 = 0;      // sum_j H_kj u_j
    double gu = 0;                      // sum_j u_j . G_kj
    double c0 = 0, c1 = 0, c2 = 0;      // sum_j q_j G_kj
    for (int j0 = 0; j0 < n; j0 += kAdjThreads) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        const bool src = j < n && is_src(st[j]);
        tp[threadIdx.x] = src ? make_double4(X[j], Y[j], Z[j], 0.0) : make_double4(0, 0, 0, 0);
        tu[threadIdx.x] = src ? U[j] : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int cnt = min(kAdjThreads, n - j0);
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
    if (k >= n) return;
    double rx = 0, ry = 0, rz = 0, rq = 0;
    if (act) {
        rx = a0 - uk.w * b0;
        ry = a1 - uk.w * b1;
        rz = a2 - uk.w * b2;
        rq = -gu;
        if (st[k] == 1) {
            const double km = K / species_mass(sp[k]);
            rq += km * (BX[k] * c0 + BY[k] * c1 + BZ[k] * c2);
        }
    }
    if (acc_x) {
        xb[k] += rx;
        yb[k] += ry;
        zb[k] += rz;
    } else {
        xb[k] = rx;
        yb[k] = ry;
        zb[k] = rz;
    }
    if (qb) qb[k] = acc_q ? qb[k] + rq : rq;
}

int force_vjp(shp_handle h, const shp_particles* p, const double* bax, const double* bay,
              const double* baz, double* xb, double* yb, double* zb, double* qb, int acc_x, int acc_q) {
    const int n = p->n;
    SHP_WS(U, double4, h, WS_ADJ_U, n);
    hipLaunchKernelGGL(k_vjp_prep, dim3(adiv(n, 256)), dim3(256), 0, h->stream, n, p->q, p->status,
                       p->species, bax, bay, baz, k_const_host(), U);
    hipLaunchKernelGGL(k_force_vjp, dim3(adiv(n, kAdjThreads)), dim3(kAdjThreads), 0, h->stream, n,
                       p->x, p->y, p->z, p->status, p->species, bax, bay, baz, U, k_const_host(),
                       kSoftening, xb, yb, zb, qb, acc_x, acc_q);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

// xh = x + v dt/2 for drifting ions (status 1, 2); w = vb1 + xb1 dt/2;
// b = dt w for targets.
__global__ void k_step_vjp_pre(int n, const double* __restrict__ X, const double* __restrict__ Y,
                               const double* __restrict__ Z, const double* __restrict__ VX,
                               const double* __restrict__ VY, const double* __restrict__ VZ,
                               const int* __restrict__ st, double dt, const double* __restrict__ XB1,
                               const double* __restrict__ YB1, const double* __restrict__ ZB1,
                               const double* __restrict__ VXB1, const double* __restrict__ VYB1,
                               const double* __restrict__ VZB1, double* __restrict__ XH,
                               double* __restrict__ YH, double* __restrict__ ZH, double* __restrict__ W,
                               double* __restrict__ B) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = is_src(st[i]);
    const double h = 0.5 * dt;
    XH[i] = mv ? X[i] + VX[i] * h : X[i];
    YH[i] = mv ? Y[i] + VY[i] * h : Y[i];
    ZH[i] = mv ? Z[i] + VZ[i] * h : Z[i];
    const double w0 = mv ? VXB1[i] + XB1[i] * h : VXB1[i];
    const double w1 = mv ? VYB1[i] + YB1[i] * h : VYB1[i];
    const double w2 = mv ? VZB1[i] + ZB1[i] * h : VZB1[i];
    W[i] = w0;
    W[n + i] = w1;
    W[2 * (size_t)n + i] = w2;
    const bool tg = st[i] == 1;
    B[i] = tg ? dt * w0 : 0.0;
    B[n + i] = tg ? dt * w1 : 0.0;
    B[2 * (size_t)n + i] = tg ? dt * w2 : 0.0;
}

// xb0 = xb1 + J^T b (in XB0 already holding J^T b); vb0 = w + xb0 dt/2.
__global__ void k_step_vjp_post(int n, const int* __restrict__ st, double dt,
                                const double* __restrict__ XB1, const double* __restrict__ YB1,
                                const double* __restrict__ ZB1, const double* __restrict__ W,
                                double* __restrict__ XB0, double* __restrict__ YB0,
                                double* __restrict__ ZB0, double* __restrict__ VXB0,
                                double* __restrict__ VYB0, double* __restrict__ VZB0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool mv = is_src(st[i]);
    const double h = mv ? 0.5 * dt : 0.0;
    const double x0 = XB1[i] + XB0[i], y0 = YB1[i] + YB0[i], z0 = ZB1[i] + ZB0[i];
    XB0[i] = x0;
    YB0[i] = y0;
    ZB0[i] = z0;
    VXB0[i] = W[i] + x0 * h;
    VYB0[i] = W[n + i] + y0 * h;
    VZB0[i] = W[2 * (size_t)n + i] + z0 * h;
}

}  // namespace
}  // namespace shp

using namespace shp;

extern "C" {

int shp_force_jvp_dev(shp_handle h, const shp_particles* p, const double* dx, const double* dy,
                      const double* dz, const double* dq, double* oax, double* oay, double* oaz) {
    SHP_REQUIRE(h && p && oax && oay && oaz, "shp_force_jvp_dev: null argument");
    SHP_REQUIRE(p->n >= 0, "shp_force_jvp_dev: n=%d", p->n);
    const int n = p->n;
    if (n == 0) return SHP_OK;
    SHP_REQUIRE(p->x && p->y && p->z && p->q && p->status && p->species,
                "shp_force_jvp_dev: null particle array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_force_jvp, dim3(adiv(n, kAdjThreads)), dim3(kAdjThreads), 0, h->stream, n,
                       p->x, p->y, p->z, p->q, p->status, p->species, dx, dy, dz, dq, k_const_host(),
                       kSoftening, oax, oay, oaz);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

int shp_force_vjp_dev(shp_handle h, const shp_particles* p, const double* bax, const double* bay,
                      const double* baz, double* xb, double* yb, double* zb, double* qb,
                      int accumulate) {
    SHP_REQUIRE(h && p && bax && bay && baz && xb && yb && zb, "shp_force_vjp_dev: null argument");
    SHP_REQUIRE(p->n >= 0, "shp_force_vjp_dev: n=%d", p->n);
    const int n = p->n;
    if (n == 0) return SHP_OK;
    SHP_REQUIRE(p->x && p->y && p->z && p->q && p->status && p->species,
                "shp_force_vjp_dev: null particle array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    return force_vjp(h, p, bax, bay, baz, xb, yb, zb, qb, accumulate, accumulate);
}

int shp_leapfrog_step_vjp_dev(shp_handle h, const shp_state* s, double dt, const double* xb1,
                              const double* yb1, const double* zb1, const double* vxb1,
                              const double* vyb1, const double* vzb1, double* xb0, double* yb0,
                              double* zb0, double* vxb0, double* vyb0, double* vzb0, double* qb) {
    SHP_REQUIRE(h && s, "shp_leapfrog_step_vjp_dev: null argument");
    SHP_REQUIRE(s->n >= 0, "shp_leapfrog_step_vjp_dev: n=%d", s->n);
    const int n = s->n;
    if (n == 0) return SHP_OK;
    SHP_REQUIRE(s->x && s->y && s->z && s->vx && s->vy && s->vz && s->q && s->status && s->species,
                "shp_leapfrog_step_vjp_dev: null state array");
    SHP_REQUIRE(xb1 && yb1 && zb1 && vxb1 && vyb1 && vzb1 && xb0 && yb0 && zb0 && vxb0 && vyb0 &&
                    vzb0 && qb,
                "shp_leapfrog_step_vjp_dev: null cotangent array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(tmp, double, h, WS_ADJ_TMP, 9 * (size_t)n);
    double *XH = tmp, *YH = tmp + n, *ZH = tmp + 2 * (size_t)n, *W = tmp + 3 * (size_t)n,
           *B = tmp + 6 * (size_t)n;
    const dim3 g(adiv(n, 256)), b(256);
    hipLaunchKernelGGL(k_step_vjp_pre, g, b, 0, h->stream, n, s->x, s->y, s->z, s->vx, s->vy, s->vz,
                       s->status, dt, xb1, yb1, zb1, vxb1, vyb1, vzb1, XH, YH, ZH, W, B);
    shp_particles p{n, XH, YH, ZH, s->q, s->status, s->species};
    // J^T b into xb0 (overwritten), charge cotangent accumulated into qb
    int rc = force_vjp(h, &p, B, B + n, B + 2 * (size_t)n, xb0, yb0, zb0, qb, 0, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_step_vjp_post, g, b, 0, h->stream, n, s->status, dt, xb1, yb1, zb1, W, xb0,
                       yb0, zb0, vxb0, vyb0, vzb0);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

}  // extern "C"
