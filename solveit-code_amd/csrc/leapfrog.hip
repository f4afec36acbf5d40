// Background E-field (pwl on a scattered triangulation) and the DKD leapfrog
// update on the FP64 device state.
//
// The field lookup replaces the reference's per-ion barycentric walk from a
// static warm start (pwl_interp_2d_scattered.c:1762-1990, called twice per ion
// by interpolation_scheme.cpp:244,247) with a quadtree point location built once
// on the host: a lane descends the tree, tests the few triangles its leaf lists
// with the reference's barycentric formula and takes the first that contains
// the point.  Inside the hull this is the triangle the walk stops in (up to
// ties on shared edges, where both give the same interpolant).  Outside the hull
// the walk is continued from the leaf's first triangle, as a walk started there
// would go; the reference's result there depends on its call history.
//
// The update kernel fuses, per ion: the kick (Coulomb + field acceleration,
// propagateVelocityNew particles.cpp:763-786), the second half drift of this
// step and the first half drift of the next (propagatePositionNew
// particles.cpp:693-739), each followed by symmetry folding and boundary
// conditions (particles.cpp:34-40, 73-177, 199-340).  It also writes the
// status masked by the next step's active range for the next force call.
#include <algorithm>
#include <climits>
#include <cmath>
#include <functional>
#include <vector>

#include "shp_internal.h"

namespace shp {

static inline int cdiv_l(long a, long b) { return (int)((a + b - 1) / b); }

void free_mesh(shp_handle h) {
    FieldMesh& m = h->mesh;
    void* ptrs[] = {m.xy, m.er, m.ez, m.tri, m.nbr, m.qnode, m.qlist};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    m = FieldMesh();
}

struct MeshView {
    const double* xy;
    const double* er;
    const double* ez;
    const int* tri;
    const int* nbr;
    const int2* qnode;
    const int* qlist;
    int ntri;
    double x0, y0, side;
    double r0, E0;
};

static MeshView view_of(const FieldMesh& m) {
    return MeshView{m.xy, m.er, m.ez, m.tri, m.nbr, m.qnode, m.qlist, m.ntri,
                    m.root[0], m.root[1], m.root[2], m.r0, m.E0};
}

// Barycentric coordinates of p in triangle t, written as in
// triangulation_search_delaunay (vertex c as base).
__device__ __forceinline__ void bary(const MeshView& M, int t, double px, double py, double& al,
                                     double& be, double& ga) {
#pragma clang fp contract(off)
    const int a = M.tri[3 * t], b = M.tri[3 * t + 1], c = M.tri[3 * t + 2];
    const double2 A = *reinterpret_cast<const double2*>(M.xy + 2 * a);
    const double2 B = *reinterpret_cast<const double2*>(M.xy + 2 * b);
    const double2 C = *reinterpret_cast<const double2*>(M.xy + 2 * c);
    const double dxa = A.x - C.x, dya = A.y - C.y;
    const double dxb = B.x - C.x, dyb = B.y - C.y;
    const double dxp = px - C.x, dyp = py - C.y;
    const double det = dxa * dyb - dya * dxb;
    al = (dxp * dyb - dyp * dxb) / det;
    be = (dxa * dyp - dya * dxp) / det;
    ga = 1.0 - al - be;
}

// Locate (px, py); returns the triangle and its barycentric coordinates.
__device__ int locate(const MeshView& M, double px, double py, double& al, double& be,
                      double& ga) {
    double x0 = M.x0, y0 = M.y0, side = M.side;
    int node = 0;
    int2 q = M.qnode[0];
    while (q.y < 0) {
        side *= 0.5;
        const int qx = px >= x0 + side;
        const int qy = py >= y0 + side;
        x0 += qx * side;
        y0 += qy * side;
        node = q.x + qx + 2 * qy;
        q = M.qnode[node];
    }
    for (int k = 0; k < q.y; k++) {
        const int t = M.qlist[q.x + k];
        bary(M, t, px, py, al, be, ga);
        if (0.0 <= al && 0.0 <= be && 0.0 <= ga) return t;
    }
    // Not inside any listed triangle: the reference walk from the leaf's first.
    int t = M.qlist[q.x];
    for (int step = 0; step <= M.ntri; step++) {
        bary(M, t, px, py, al, be, ga);
        if (0.0 <= al && 0.0 <= be && 0.0 <= ga) break;
        const int* nb = M.nbr + 3 * t;
        if (al < 0.0 && 0 <= nb[0]) { t = nb[0]; continue; }
        if (be < 0.0 && 0 <= nb[1]) { t = nb[1]; continue; }
        if (ga < 0.0 && 0 <= nb[2]) { t = nb[2]; continue; }
        break;  // outside the hull: extrapolate from this hull triangle
    }
    return t;
}

// FP64 field at (x, y, z) in V/m; returns the triangle (-1 when gated off).
__device__ __forceinline__ int efield_f64(const MeshView& M, double x, double y, double z,
                                          double& ex, double& ey, double& ez) {
    const double rho = sqrt(x * x + y * y);
    const double zn = z / M.r0;
    ex = ey = ez = 0.0;
    if (zn < 0.0 || zn > 140.0) return -1;
    double al, be, ga;
    const int t = locate(M, rho / M.r0, zn, al, be, ga);
    const int a = M.tri[3 * t], b = M.tri[3 * t + 1], c = M.tri[3 * t + 2];
    const double vr = al * M.er[a] + be * M.er[b] + ga * M.er[c];
    const double vz = al * M.ez[a] + be * M.ez[b] + ga * M.ez[c];
    double cs = 1.0, sn = 0.0;
    if (rho > 0.0) {
        cs = x / rho;
        sn = y / rho;
    }
    ex = M.E0 * vr * cs;
    ey = M.E0 * vr * sn;
    ez = M.E0 * vz;
    return t;
}

__global__ void k_efield(MeshView M, int n, const double* __restrict__ X,
                         const double* __restrict__ Y, const double* __restrict__ Z,
                         double* __restrict__ ex, double* __restrict__ ey, double* __restrict__ ez,
                         int* __restrict__ tri_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a, b, c;
    const int t = efield_f64(M, X[i], Y[i], Z[i], a, b, c);
    ex[i] = a;
    ey[i] = b;
    ez[i] = c;
    if (tri_out) tri_out[i] = t;
}

// ------------------------------------------------------------------ energies ----
// FP64 restatement of the reference's energy bookkeeping (particles.cpp:535-638,
// integration_schemes.cpp:799-835); the particle constants are the float ones
// of constants.h:28-47 widened to double.

constexpr double kElemChargeD = (double)1.6021766e-19f;
constexpr double kAmuD = (double)1.66053904e-27f;
constexpr int kEnergyThreads = 256;

__device__ __forceinline__ bool present(int s) { return s == 1 || s == 2; }

// Per-block partial sums of 0.5 amu m v^2.
__global__ __launch_bounds__(kEnergyThreads) void k_energy_ke(
    int n, const double* __restrict__ vx, const double* __restrict__ vy,
    const double* __restrict__ vz, const int* __restrict__ st, const int* __restrict__ sp,
    double* __restrict__ part) {
    __shared__ double red[kEnergyThreads];
    double a = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (present(st[i]))
            a += 0.5 * kAmuD * species_mass(sp[i]) * (vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i]);
    red[threadIdx.x] = a;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Per-block partial sums of q_i sum_{j != i} q_j / sqrt(r^2 + eps^2) over
// LDS-staged source tiles (FP64).
__global__ __launch_bounds__(kEnergyThreads) void k_energy_pe(
    int n, const double* __restrict__ X, const double* __restrict__ Y,
    const double* __restrict__ Z, const double* __restrict__ Q, const int* __restrict__ st,
    double eps2, double* __restrict__ part) {
    __shared__ double4 tile[kEnergyThreads];
    __shared__ double red[kEnergyThreads];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool tgt = i < n && present(st[i]);
    const double xi = tgt ? X[i] : 0.0, yi = tgt ? Y[i] : 0.0, zi = tgt ? Z[i] : 0.0;
    double a = 0.0;
    for (int j0 = 0; j0 < n; j0 += kEnergyThreads) {
        __syncthreads();
        const int j = j0 + threadIdx.x;
        tile[threadIdx.x] = j < n && present(st[j]) ? make_double4(X[j], Y[j], Z[j], Q[j])
                                                   : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int cnt = min(kEnergyThreads, n - j0);
#pragma unroll 4
        for (int k = 0; k < cnt; k++) {
            const double4 p = tile[k];
            const double dx = xi - p.x, dy = yi - p.y, dz = zi - p.z;
            const double w = p.w * rsqrt_f64(fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2))));
            a += j0 + k == i ? 0.0 : w;
        }
    }
    red[threadIdx.x] = tgt ? Q[i] * a : 0.0;
    __syncthreads();
    for (int s = kEnergyThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Field energy along the reference's z-then-radial path (left Riemann sums).
__global__ void k_field_energy(MeshView M, int n, const double* __restrict__ X0,
                               const double* __restrict__ Y0, const double* __restrict__ Z0,
                               const double* __restrict__ X, const double* __restrict__ Y,
                               const double* __restrict__ Z, const double* __restrict__ Q,
                               const int* __restrict__ ns_arr, int nsteps, double* __restrict__ fe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int ns = ns_arr ? ns_arr[i] : nsteps;
    const double x0 = X0[i], y0 = Y0[i], z0 = Z0[i];
    const double dx = X[i] - x0, dy = Y[i] - y0, dz = Z[i] - z0;
    double fz = 0.0, fr = 0.0, ex, ey, ez;
    for (int k = 0; k < ns; k++) {
        efield_f64(M, x0, y0, z0 + k * dz / ns, ex, ey, ez);
        fz += ez * dz / ns;
    }
    const double zt = z0 + dz;
    for (int k = 0; k < ns; k++) {
        efield_f64(M, x0 + k * dx / ns, y0 + k * dy / ns, zt, ex, ey, ez);
        fr += ex * dx / ns + ey * dy / ns;
    }
    fe[i] += (fz + fr) * (-kElemChargeD * Q[i]);
}

// ------------------------------------------------------------ update kernels ----

struct UpdateParams {
    int sym;            // SHP_SYM_*
    int nsec;
    double dom[3];      // translational cell
    int axes[3];        // wrap axes (domain > 0)
    double bc[4];       // |x|,|y|,|z|,rho limits
    double bc_time;
    double e_over_amu;  // ELEM_CHARGE / AMU (particles.cpp:500)
    int use_efield;
    // Outputs of the resident runs (null otherwise): the background field of
    // each kick, [3][n], and the boundary-condition event queue.
    float* e2;
    unsigned int* ev;
    unsigned long long* evkey;
    int* evn;
    int evcap;
    long long seq;  // number of the kernel's first drift since the upload
};

// One boundary-condition event in writeToEventsFile's layout, queued with the
// key (2 drift + pass) << 32 | index: the reference enforces the spatial
// limits over all ions, then the temporal one (particles.cpp:73-92), each in
// index order, after every drift.
__device__ void bc_event(const UpdateParams& P, long long seq, int pass, double t, double x, double y,
                         double z, double vx, double vy, double vz, int i, int sp) {
    const int k = atomicAdd(P.evn, 1);
    if (k >= P.evcap) return;  // cannot happen: a status-1 ion leaves once per upload
    unsigned int* r = P.ev + (size_t)k * kEventWords;
    r[0] = __float_as_uint((float)t);
    r[1] = __float_as_uint((float)x);
    r[2] = __float_as_uint((float)y);
    r[3] = __float_as_uint((float)z);
    r[4] = __float_as_uint((float)vx);
    r[5] = __float_as_uint((float)vy);
    r[6] = __float_as_uint((float)vz);
    r[7] = (unsigned int)i;
    r[8] = (unsigned int)sp;
    r[9] = 0u;  // ParticleStatus::outOfBounds
    P.evkey[k] = ((unsigned long long)(2 * seq + pass) << 32) | (unsigned int)i;
}

__device__ __forceinline__ void fold_and_bound(const UpdateParams& P, double& x, double& y,
                                               double& z, double& vx, double& vy, double vz, int& st,
                                               double t_now, double t_inj, int i, int sp, long long seq) {
    if (P.sym == SHP_SYM_ROTATIONAL) {
        if (P.nsec == 4) {
            // rotateToFirstQuadrant, particles.cpp:232-262
            if (x < 0 && y < 0) {
                x = -x; y = -y; vx = -vx; vy = -vy;
            } else if (x < 0) {
                double tx = y, ty = -x; x = tx; y = ty;
                tx = vy; ty = -vx; vx = tx; vy = ty;
            } else if (y < 0) {
                double tx = -y, ty = x; x = tx; y = ty;
                tx = -vy; ty = vx; vx = tx; vy = ty;
            }
        } else if (P.nsec > 1) {
            const double w = 2.0 * M_PI / P.nsec;
            double phi = atan2(y, x);
            if (phi < 0) phi += 2.0 * M_PI;
            const int k = (int)floor(phi / w);
            if (k > 0) {
                const double c = cos(-k * w), s = sin(-k * w);
                double tx = c * x - s * y, ty = s * x + c * y;
                x = tx; y = ty;
                tx = c * vx - s * vy; ty = s * vx + c * vy;
                vx = tx; vy = ty;
            }
        }
    } else if (P.sym == SHP_SYM_TRANSLATIONAL) {
        // translateToSymmetryDomain, particles.cpp:282-330
        double* p[3] = {&x, &y, &z};
        for (int c = 0; c < 3; c++) {
            if (!P.axes[c]) continue;
            const double D = P.dom[c];
            double t = *p[c] + D / 2;
            if (t >= 0) {
                *p[c] = fmod(t, D) - D / 2;
            } else {
                t = fmod(t, D);
                *p[c] = t < 0 ? t + D / 2 : t - D / 2;
            }
        }
    }
    if (st == 1) {
        bool out = false;
        if (P.bc[0] != 0 && fabs(x) > P.bc[0]) out = true;
        if (P.bc[1] != 0 && fabs(y) > P.bc[1]) out = true;
        if (P.bc[2] != 0 && fabs(z) > P.bc[2]) out = true;
        if (P.bc[3] != 0 && sqrt(x * x + y * y) > P.bc[3]) out = true;
        const bool late = P.bc_time != 0 && t_now - t_inj > P.bc_time;
        if (out || late) {
            st = 0;
            if (P.evn) bc_event(P, seq, out ? 0 : 1, t_now, x, y, z, vx, vy, vz, i, sp);
        }
    }
}

// One half drift of ion state over [t_start, t_start + h].
__device__ __forceinline__ void half_drift(const UpdateParams& P, double& x, double& y, double& z,
                                           double& vx, double& vy, double vz, int& st,
                                           double t_inj, double t_start, double h, int i, int sp,
                                           long long seq) {
    const double t_end = t_start + h;
    if (t_inj > t_end) return;  // not injected yet: outside [first, last)
    if (st == 1 || st == 2) {
        const double s = fmin(fmax(t_end - t_inj, 0.0), h);
        x += vx * s;
        y += vy * s;
        z += vz * s;
    }
    fold_and_bound(P, x, y, z, vx, vy, vz, st, t_end, t_inj, i, sp, seq);
}

// First half drift of a run; writes the active-masked status for the force.
__global__ void k_drift(UpdateParams P, int n, double* __restrict__ X, double* __restrict__ Y,
                        double* __restrict__ Z, double* __restrict__ VX, double* __restrict__ VY,
                        const double* __restrict__ VZ, int* __restrict__ status,
                        const int* __restrict__ species, const double* __restrict__ tinj, double t_start,
                        double h, int* __restrict__ eff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double x = X[i], y = Y[i], z = Z[i], vx = VX[i], vy = VY[i];
    int st = status[i];
    const double ti = tinj[i];
    half_drift(P, x, y, z, vx, vy, VZ[i], st, ti, t_start, h, i, species[i], P.seq);
    X[i] = x; Y[i] = y; Z[i] = z; VX[i] = vx; VY[i] = vy;
    status[i] = st;
    eff[i] = ti <= t_start + h ? st : 0;
}

// Kick over [t, t+dt] (ions active since the drift to t + dt/2), then the
// second half drift of this step and, when `next` is set, the first half drift
// of the next step.
__global__ void k_kick_drift(UpdateParams P, MeshView M, int n, double* __restrict__ X,
                             double* __restrict__ Y, double* __restrict__ Z,
                             double* __restrict__ VX, double* __restrict__ VY,
                             double* __restrict__ VZ, const double* __restrict__ Q,
                             int* __restrict__ status, const int* __restrict__ species,
                             const double* __restrict__ tinj, const double* __restrict__ AX,
                             const double* __restrict__ AY, const double* __restrict__ AZ,
                             double t, double dt, int next, int* __restrict__ eff,
                             const double* __restrict__ tdev) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (tdev) t = *tdev;  // time held in device memory (graph replays)
    double x = X[i], y = Y[i], z = Z[i];
    double vx = VX[i], vy = VY[i], vz = VZ[i];
    int st = status[i];
    const double ti = tinj[i];
    const double h = 0.5 * dt;
    if (st == 1 && ti <= t + h) {
        double ax = AX[i], ay = AY[i], az = AZ[i];
        double ex = 0.0, ey = 0.0, ez = 0.0;
        if (P.use_efield) {
            efield_f64(M, x, y, z, ex, ey, ez);
            const double f = P.e_over_amu * Q[i] / species_mass(species[i]);
            ax += f * ex;
            ay += f * ey;
            az += f * ez;
        }
        if (P.e2) {  // Efield2 of the kicked ion (particles.cpp:512)
            P.e2[i] = (float)ex;
            P.e2[n + i] = (float)ey;
            P.e2[2 * (size_t)n + i] = (float)ez;
        }
        const double s = fmin(fmax(t + dt - ti, 0.0), dt);
        vx += ax * s;
        vy += ay * s;
        vz += az * s;
    }
    const int sp = species[i];
    half_drift(P, x, y, z, vx, vy, vz, st, ti, t + h, h, i, sp, P.seq);
    if (next) half_drift(P, x, y, z, vx, vy, vz, st, ti, t + dt, h, i, sp, P.seq + 1);
    X[i] = x; Y[i] = y; Z[i] = z;
    VX[i] = vx; VY[i] = vy; VZ[i] = vz;
    status[i] = st;
    if (next) eff[i] = ti <= t + dt + h ? st : 0;
}

static UpdateParams make_params(const shp_step_opts* o) {
    UpdateParams P{};
    P.sym = o->sym.type;
    P.nsec = o->sym.nsec;
    for (int c = 0; c < 3; c++) {
        P.dom[c] = o->sym.domain[c];
        P.axes[c] = o->sym.domain[c] > 0;
    }
    for (int c = 0; c < 4; c++) P.bc[c] = o->bc_space[c];
    P.bc_time = o->bc_time;
    P.e_over_amu = (double)1.6021766e-19f / (double)1.66053904e-27f;
    P.use_efield = o->use_efield;
    return P;
}

// The resident run's outputs (field of each kick, event queue) while
// shp_state_run* drives the leapfrog; P.seq is set per launch.
static void attach_outputs(shp_handle h, UpdateParams& P) {
    if (!h->res_events_on) return;
    P.e2 = h->res_e2;
    P.ev = h->res_ev;
    P.evkey = h->res_evkey;
    P.evn = h->res_evn;
    P.evcap = h->res_cap;
}

// Accuracy metric of Integrator::compareAcceleration
// (integration_schemes.cpp:374-470) in FP64.  Per ion with status 1:
// rel = |a - a_ref|^2 / |a_ref|^2, 0 when both are 0.  Each block folds its
// grid-strided ions in index order per lane, then a fixed tree; the maximum
// keeps the smallest index among equal values (the reference's first index).
constexpr int kCmpBlocks = 1024;
__global__ __launch_bounds__(256) void k_cmp_partial(int n, const int* __restrict__ st,
                                                     const double* __restrict__ ax,
                                                     const double* __restrict__ ay,
                                                     const double* __restrict__ az,
                                                     const double* __restrict__ rx,
                                                     const double* __restrict__ ry,
                                                     const double* __restrict__ rz,
                                                     double* __restrict__ psum, double* __restrict__ pmax,
                                                     int* __restrict__ pidx, int* __restrict__ pcnt) {
    __shared__ double ss[256], sm[256];
    __shared__ int si[256], sc[256];
    double sum = 0.0, mx = 0.0;
    int idx = INT_MAX, cnt = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        if (st[i] != 1) continue;
        cnt++;
        const double dx = ax[i] - rx[i], dy = ay[i] - ry[i], dz = az[i] - rz[i];
        const double d = dx * dx + dy * dy + dz * dz;
        const double nr = rx[i] * rx[i] + ry[i] * ry[i] + rz[i] * rz[i];
        const double rel = (d == 0.0 && nr == 0.0) ? 0.0 : d / nr;
        sum += rel;
        if (rel > mx) {
            mx = rel;
            idx = i;
        }
    }
    const int t = threadIdx.x;
    ss[t] = sum;
    sm[t] = mx;
    si[t] = idx;
    sc[t] = cnt;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            ss[t] += ss[t + w];
            sc[t] += sc[t + w];
            if (sm[t + w] > sm[t] || (sm[t + w] == sm[t] && si[t + w] < si[t])) {
                sm[t] = sm[t + w];
                si[t] = si[t + w];
            }
        }
        __syncthreads();
    }
    if (t == 0) {
        psum[blockIdx.x] = ss[0];
        pmax[blockIdx.x] = sm[0];
        pidx[blockIdx.x] = si[0];
        pcnt[blockIdx.x] = sc[0];
    }
}

__global__ void k_cmp_final(int nblk, const double* __restrict__ psum, const double* __restrict__ pmax,
                            const int* __restrict__ pidx, const int* __restrict__ pcnt,
                            double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double sum = 0.0, mx = 0.0;
    int idx = INT_MAX;
    long cnt = 0;
    for (int b = 0; b < nblk; b++) {
        sum += psum[b];
        cnt += pcnt[b];
        if (pmax[b] > mx || (pmax[b] == mx && pidx[b] < idx)) {
            mx = pmax[b];
            idx = pidx[b];
        }
    }
    out[0] = (double)cnt;
    out[1] = sqrt(sum / (double)cnt);  // NaN without valid ions, as in the reference
    out[2] = sqrt(mx);
    out[3] = idx == INT_MAX ? 0.0 : (double)idx;
}

}  // namespace shp

namespace shp {

__global__ void k_time_advance(double* __restrict__ t, double dt) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *t += dt;
}

// First half drift of a run; writes the effective (injected) status.
int lf_begin(shp_handle h, const shp_state* s, const shp_step_opts* o, double t0, double dt, int* eff) {
    const UpdateParams P = make_params(o);
    const int n = s->n;
    hipLaunchKernelGGL(k_drift, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, P, n, s->x, s->y, s->z,
                       s->vx, s->vy, s->vz, s->status, s->species, s->t_inj, t0, 0.5 * dt, eff);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

// Kick, second half drift and (next != 0) the next first half drift; the
// step's start time is *tdev when tdev is given (then advanced by dt).
int lf_kick_drift(shp_handle h, const shp_state* s, const shp_step_opts* o, double t, const double* tdev,
                  double dt, int next, const double* ax, const double* ay, const double* az, int* eff) {
    SHP_REQUIRE(!o->use_efield || h->mesh.ready, "leapfrog: E-field requested but no mesh");
    const UpdateParams P = make_params(o);
    const MeshView M = h->mesh.ready ? view_of(h->mesh) : MeshView{};
    const int n = s->n;
    hipLaunchKernelGGL(k_kick_drift, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, P, M, n, s->x, s->y,
                       s->z, s->vx, s->vy, s->vz, s->q, s->status, s->species, s->t_inj, ax, ay, az, t,
                       dt, next, eff, tdev);
    if (tdev) hipLaunchKernelGGL(k_time_advance, dim3(1), dim3(64), 0, h->stream, (double*)tdev, dt);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}


// Resident state from the reference's float AoS arrays: FP64 positions,
// velocities, charge, injection time; status from acceleration.w, species
// from velocity.w; the acceleration starts at the array's xyz.
__global__ void k_res_from_aos(const float4* __restrict__ pos4, const float4* __restrict__ vel4,
                               const float4* __restrict__ acc4, const float* __restrict__ tinj, int n,
                               double* x, double* y, double* z, double* vx, double* vy, double* vz,
                               double* q, double* ti, int* status, int* species, double* ax) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 p = pos4[i], v = vel4[i], a = acc4[i];
    x[i] = p.x;
    y[i] = p.y;
    z[i] = p.z;
    q[i] = p.w;
    vx[i] = v.x;
    vy[i] = v.y;
    vz[i] = v.z;
    ti[i] = tinj[i];
    status[i] = a.w == 1.0f ? 1 : (a.w == 2.0f ? 2 : (a.w == 0.0f ? 0 : -1));
    species[i] = (int)v.w;
    ax[i] = a.x;
    ax[n + i] = a.y;
    ax[2 * (size_t)n + i] = a.z;
}

// Back to float AoS: position (x, y, z, q), velocity (vx, vy, vz, species),
// acceleration (the last Coulomb acceleration, status).
__global__ void k_res_to_aos(const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const double* __restrict__ vx,
                             const double* __restrict__ vy, const double* __restrict__ vz,
                             const double* __restrict__ ax, const double* __restrict__ ay,
                             const double* __restrict__ az, const int* __restrict__ status, int n,
                             float4* __restrict__ pos4, float4* __restrict__ vel4, float4* __restrict__ acc4) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pos4[i] = make_float4((float)x[i], (float)y[i], (float)z[i], pos4[i].w);
    vel4[i] = make_float4((float)vx[i], (float)vy[i], (float)vz[i], vel4[i].w);
    acc4[i] = make_float4((float)ax[i], (float)ay[i], (float)az[i], (float)status[i]);
}


// ------------------------------------------------------------ substeps ----
// LeapfrogDKDSubsteps (integration_schemes.cpp:899-1529): levels 0..M with
// dt_l = dt_{l-1} / 3, a level per ion; all ions drift together by dt_M / 2,
// and each finest kick kicks the ions whose level is at least the kick's top
// level by their own level's dt (variableTimestepKick :1216-1250).

constexpr int kMaxSubLevels = 8;

struct SubstepTable {
    int M;
    int type[kMaxSubLevels];        // BoundaryConditionOptions of each level
    double bc[kMaxSubLevels][4];    // |x|, |y|, |z|, rho limits
    double bt[kMaxSubLevels];       // time-since-injection limit
    double dtl[kMaxSubLevels];      // dt of each level
};

// The finest level (from M down to `level`) whose limit admits v, extended
// down over the levels sharing that limit; -1 when none does
// (checkSpatialConditions :944-1066, checkTemporalConditions :1073-1110).
__device__ __forceinline__ int ss_scan(const SubstepTable& T, int level, int axis, double v) {
    int possible = -1;
    double bound = 0.0;
    for (int i = T.M; i >= level; i--) {
        const double bc = axis < 4 ? T.bc[i][axis] : T.bt[i];
        if (possible == -1) {
            if (v < bc || bc == 0) {
                possible = i;
                bound = bc;
            }
        } else {
            if (bound == bc) possible--;
            else break;
        }
    }
    return possible;
}

// checkConditions(index, level) (:919-937): the ion's new level.
__device__ int ss_level_of(const SubstepTable& T, int level, int cur, double x, double y, double z,
                           double tnow, double tinj) {
    const int type = T.type[level + 1];
    if (type == 0) return cur;
    int sp = cur;
    if (type == 1 || type == 3) {
        sp = ss_scan(T, level, 0, fabs(x));
        sp = max(sp, ss_scan(T, level, 1, fabs(y)));
        sp = max(sp, ss_scan(T, level, 2, fabs(z)));
        sp = max(sp, ss_scan(T, level, 3, sqrt(x * x + y * y)));
        if (type == 1) return sp;
    }
    const int tp = ss_scan(T, level, 4, tnow - tinj);
    return type == 2 ? tp : max(tp, sp);
}

// setParticleStatus (:1193-1211) over the injected ions, and the status the
// force sees (0 for ions not injected yet).
__global__ void k_ss_status(int n, const double* __restrict__ tinj, double tpos, const int* __restrict__ level,
                            int top, int* __restrict__ status, int* __restrict__ eff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool injected = tinj[i] <= tpos;
    int st = status[i];
    if (injected && st != 0 && st != -1) {
        st = level[i] >= top ? 1 : 2;
        status[i] = st;
    }
    eff[i] = injected ? st : 0;
}

// The kick of the status-1 ions: Coulomb plus field acceleration over the
// ion's own level dt, clamped from its injection time (window ending at
// tvel + dt_level).
__global__ void k_ss_kick(UpdateParams P, MeshView M, SubstepTable T, int n, const double* __restrict__ X,
                          const double* __restrict__ Y, const double* __restrict__ Z, double* __restrict__ VX,
                          double* __restrict__ VY, double* __restrict__ VZ, const double* __restrict__ Q,
                          const int* __restrict__ species, const double* __restrict__ tinj,
                          const int* __restrict__ eff, const int* __restrict__ level,
                          const double* __restrict__ AX, const double* __restrict__ AY,
                          const double* __restrict__ AZ, double tvel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || eff[i] != 1) return;
    double ax = AX[i], ay = AY[i], az = AZ[i];
    double ex = 0.0, ey = 0.0, ez = 0.0;
    if (P.use_efield) {
        efield_f64(M, X[i], Y[i], Z[i], ex, ey, ez);
        const double f = P.e_over_amu * Q[i] / species_mass(species[i]);
        ax += f * ex;
        ay += f * ey;
        az += f * ez;
    }
    if (P.e2) {  // Efield2 of the kicked ion (particles.cpp:512)
        P.e2[i] = (float)ex;
        P.e2[n + i] = (float)ey;
        P.e2[2 * (size_t)n + i] = (float)ez;
    }
    const double dl = T.dtl[level[i]];
    const double s = fmin(fmax(tvel + dl - tinj[i], 0.0), dl);
    VX[i] += ax * s;
    VY[i] += ay * s;
    VZ[i] += az * s;
}

__global__ void k_ss_fill(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// crossoverParticles(L) (:1117-1127) over the injected ions.
__global__ void k_ss_crossover(SubstepTable T, int L, int n, const double* __restrict__ X,
                               const double* __restrict__ Y, const double* __restrict__ Z,
                               const double* __restrict__ tinj, double tpos, int* __restrict__ level) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || tinj[i] > tpos) return;
    const int lv = level[i];
    if (lv == L || lv == L + 1) level[i] = ss_level_of(T, L, lv, X[i], Y[i], Z[i], tpos, tinj[i]);
}

}  // namespace shp

using namespace shp;

extern "C" {

int shp_efield_setup(shp_handle h, const shp_field_mesh* m) {
    SHP_REQUIRE(h && m && m->xy && m->er && m->ez && m->tri && m->nbr, "shp_efield_setup: null");
    SHP_REQUIRE(m->nnodes >= 3 && m->ntri >= 1, "shp_efield_setup: %d nodes, %d triangles",
                m->nnodes, m->ntri);
    SHP_REQUIRE(m->r0 > 0, "shp_efield_setup: r0 must be positive");
    for (int k = 0; k < 3 * m->ntri; k++)
        SHP_REQUIRE(m->tri[k] >= 0 && m->tri[k] < m->nnodes && m->nbr[k] < m->ntri,
                    "shp_efield_setup: triangle/neighbour index out of range at %d", k);
    SHP_HIP_CHECK(hipSetDevice(h->device));
    // Quadtree over the bounding square of the nodes.
    double x0 = m->xy[0], x1 = x0, y0 = m->xy[1], y1 = y0;
    for (int i = 0; i < m->nnodes; i++) {
        x0 = std::min(x0, m->xy[2 * i]);
        x1 = std::max(x1, m->xy[2 * i]);
        y0 = std::min(y0, m->xy[2 * i + 1]);
        y1 = std::max(y1, m->xy[2 * i + 1]);
    }
    double side = std::max(x1 - x0, y1 - y0) * (1.0 + 1e-9) + 1e-300;
    struct TB { double x0, x1, y0, y1; };
    std::vector<TB> tb(m->ntri);
    for (int t = 0; t < m->ntri; t++) {
        TB b{1e300, -1e300, 1e300, -1e300};
        for (int k = 0; k < 3; k++) {
            const int v = m->tri[3 * t + k];
            b.x0 = std::min(b.x0, m->xy[2 * v]);
            b.x1 = std::max(b.x1, m->xy[2 * v]);
            b.y0 = std::min(b.y0, m->xy[2 * v + 1]);
            b.y1 = std::max(b.y1, m->xy[2 * v + 1]);
        }
        tb[t] = b;
    }
    std::vector<int2> nodes(1);
    std::vector<int> list;
    const int kLeafMax = 8, kDepthMax = 20;
    std::function<void(int, double, double, double, std::vector<int>&, int)> build =
        [&](int node, double cx, double cy, double s, std::vector<int>& ts, int depth) {
            if ((int)ts.size() <= kLeafMax || depth >= kDepthMax) {
                int off = (int)list.size();
                if (ts.empty()) {
                    // nearest triangle (by centroid) to the cell centre
                    double mx = cx + s / 2, my = cy + s / 2, best = 1e300;
                    int bt = 0;
                    for (int t = 0; t < m->ntri; t++) {
                        double gx = (tb[t].x0 + tb[t].x1) / 2 - mx, gy = (tb[t].y0 + tb[t].y1) / 2 - my;
                        double d = gx * gx + gy * gy;
                        if (d < best) { best = d; bt = t; }
                    }
                    list.push_back(bt);
                    nodes[node] = make_int2(off, 0);
                } else {
                    list.insert(list.end(), ts.begin(), ts.end());
                    nodes[node] = make_int2(off, (int)ts.size());
                }
                return;
            }
            int first = (int)nodes.size();
            nodes.resize(nodes.size() + 4);
            nodes[node] = make_int2(first, -1);
            double hs = s / 2;
            for (int q = 0; q < 4; q++) {
                double qx0 = cx + (q & 1) * hs, qy0 = cy + (q >> 1) * hs;
                std::vector<int> sub;
                for (int t : ts)
                    if (tb[t].x1 >= qx0 && tb[t].x0 <= qx0 + hs && tb[t].y1 >= qy0 && tb[t].y0 <= qy0 + hs)
                        sub.push_back(t);
                build(first + q, qx0, qy0, hs, sub, depth + 1);
            }
        };
    {
        std::vector<int> all(m->ntri);
        for (int t = 0; t < m->ntri; t++) all[t] = t;
        build(0, x0, y0, side, all, 0);
    }
    // A leaf that lists no triangle must still hand the walk a start: make its
    // count 0 but keep its one-entry list (locate() then walks from it).
    free_mesh(h);
    FieldMesh& M = h->mesh;
    auto up = [&](void** dst, const void* src, size_t bytes) -> int {
        SHP_HIP_CHECK(hipMalloc(dst, bytes));
        SHP_HIP_CHECK(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
        return SHP_OK;
    };
    int rc = 0;
    rc |= up((void**)&M.xy, m->xy, sizeof(double) * 2 * m->nnodes);
    rc |= up((void**)&M.er, m->er, sizeof(double) * m->nnodes);
    rc |= up((void**)&M.ez, m->ez, sizeof(double) * m->nnodes);
    rc |= up((void**)&M.tri, m->tri, sizeof(int) * 3 * m->ntri);
    rc |= up((void**)&M.nbr, m->nbr, sizeof(int) * 3 * m->ntri);
    rc |= up((void**)&M.qnode, nodes.data(), sizeof(int2) * nodes.size());
    rc |= up((void**)&M.qlist, list.data(), sizeof(int) * list.size());
    if (rc) return SHP_ERR_HIP;
    M.nnodes = m->nnodes;
    M.ntri = m->ntri;
    M.nqnodes = (int)nodes.size();
    M.nqlist = (int)list.size();
    M.root[0] = x0;
    M.root[1] = y0;
    M.root[2] = side;
    M.r0 = m->r0;
    M.E0 = m->E0;
    M.ready = true;
    return SHP_OK;
}

int shp_efield_dev(shp_handle h, int n, const double* x, const double* y, const double* z,
                   double* ex, double* ey, double* ez, int* tri_out) {
    SHP_REQUIRE(h && x && y && z && ex && ey && ez, "shp_efield_dev: null argument");
    SHP_REQUIRE(h->mesh.ready, "shp_efield_dev: no field mesh (call shp_efield_setup)");
    SHP_REQUIRE(n >= 0, "shp_efield_dev: n=%d", n);
    if (n == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_efield, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, view_of(h->mesh), n,
                       x, y, z, ex, ey, ez, tri_out);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

int shp_leapfrog_run(shp_handle h, shp_state* s, const shp_step_opts* o, double t0, double dt,
                     int nsteps, double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && s && o && ax && ay && az, "shp_leapfrog_run: null argument");
    SHP_REQUIRE(s->n >= 0 && nsteps >= 0 && dt >= 0, "shp_leapfrog_run: n=%d steps=%d dt=%g",
                s->n, nsteps, dt);
    SHP_REQUIRE(s->x && s->y && s->z && s->vx && s->vy && s->vz && s->q && s->status &&
                    s->species && s->t_inj,
                "shp_leapfrog_run: null state array");
    SHP_REQUIRE(!o->use_efield || h->mesh.ready, "shp_leapfrog_run: E-field requested but no mesh");
    SHP_REQUIRE(o->method == SHP_METHOD_DIRECT || o->method == SHP_METHOD_FMM ||
                    o->method == SHP_METHOD_BH,
                "shp_leapfrog_run: method %d not available", o->method);
    if (s->n == 0 || nsteps == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    const int n = s->n;
    SHP_WS(eff, int, h, WS_EFF_STATUS, n);
    UpdateParams P = make_params(o);
    attach_outputs(h, P);
    const MeshView M = h->mesh.ready ? view_of(h->mesh) : MeshView{};
    const dim3 grid(cdiv_l(n, 256)), block(256);
    P.seq = h->drift_seq++;
    hipLaunchKernelGGL(k_drift, grid, block, 0, h->stream, P, n, s->x, s->y, s->z, s->vx, s->vy,
                       s->vz, s->status, s->species, s->t_inj, t0, 0.5 * dt, eff);
    SHP_HIP_CHECK(hipGetLastError());
    for (int k = 0; k < nsteps; k++) {
        const double t = t0 + k * dt;
        shp_particles p{n, s->x, s->y, s->z, s->q, eff, s->species};
        int rc = o->method == SHP_METHOD_FMM
                     ? shp_fmm_sym_dev(h, &p, &o->sym, o->fmm_order, o->leaf_level, o->leaf_size,
                                       o->fmm_ring > 0 ? o->fmm_ring : 2, ax, ay, az)
                 : o->method == SHP_METHOD_BH
                     ? shp_bh_dev(h, &p, (float)o->theta, &o->sym, ax, ay, az)
                     : shp_direct_dev(h, &p, o->precision, &o->sym, ax, ay, az);
        if (rc) return rc;
        SHP_HIP_CHECK(hipEventRecord(h->ev2, h->stream));
        P.seq = h->drift_seq;
        h->drift_seq += k + 1 < nsteps ? 2 : 1;
        hipLaunchKernelGGL(k_kick_drift, grid, block, 0, h->stream, P, M, n, s->x, s->y, s->z, s->vx,
                           s->vy, s->vz, s->q, s->status, s->species, s->t_inj, ax, ay, az, t, dt,
                           k + 1 < nsteps ? 1 : 0, eff, nullptr);
        SHP_HIP_CHECK(hipGetLastError());
        SHP_HIP_CHECK(hipEventRecord(h->ev3, h->stream));
        h->timed_update = true;
    }
    return SHP_OK;
}

int shp_energy_dev(shp_handle h, const shp_particles* p, const double* vx, const double* vy,
                   const double* vz, double* out) {
    SHP_REQUIRE(h && p && out, "shp_energy_dev: null argument");
    SHP_REQUIRE(p->n >= 0, "shp_energy_dev: n=%d", p->n);
    out[0] = out[1] = 0.0;
    const int n = p->n;
    if (n == 0) return SHP_OK;
    SHP_REQUIRE(p->x && p->y && p->z && p->q && p->status && p->species && vx && vy && vz,
                "shp_energy_dev: null array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    const int nb_ke = std::min(cdiv_l(n, kEnergyThreads), 1024);
    const int nb_pe = cdiv_l(n, kEnergyThreads);
    SHP_WS(part, double, h, WS_PARTIAL, (size_t)nb_ke + nb_pe);
    hipLaunchKernelGGL(k_energy_ke, dim3(nb_ke), dim3(kEnergyThreads), 0, h->stream, n, vx, vy, vz,
                       p->status, p->species, part);
    hipLaunchKernelGGL(k_energy_pe, dim3(nb_pe), dim3(kEnergyThreads), 0, h->stream, n, p->x, p->y,
                       p->z, p->q, p->status, kSoftening, part + nb_ke);
    SHP_HIP_CHECK(hipGetLastError());
    std::vector<double> hp((size_t)nb_ke + nb_pe);
    SHP_HIP_CHECK(hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost,
                                 h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    long double ke = 0.0L, pe = 0.0L;
    for (int b = 0; b < nb_ke; b++) ke += hp[b];
    for (int b = 0; b < nb_pe; b++) pe += hp[nb_ke + b];
    out[0] = (double)ke;
    out[1] = (double)(0.5L * pe * (long double)(k_const_host() * kAmuD));
    return SHP_OK;
}

int shp_field_energy_dev(shp_handle h, int n, const double* x0, const double* y0, const double* z0,
                         const double* x, const double* y, const double* z, const double* q,
                         const int* nsteps_arr, int nsteps, double* fe) {
    SHP_REQUIRE(h && fe, "shp_field_energy_dev: null argument");
    SHP_REQUIRE(n >= 0, "shp_field_energy_dev: n=%d", n);
    SHP_REQUIRE(h->mesh.ready, "shp_field_energy_dev: no field mesh (shp_efield_setup)");
    SHP_REQUIRE(nsteps_arr || nsteps >= 1, "shp_field_energy_dev: nsteps=%d", nsteps);
    if (n == 0) return SHP_OK;
    SHP_REQUIRE(x0 && y0 && z0 && x && y && z && q, "shp_field_energy_dev: null array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    hipLaunchKernelGGL(k_field_energy, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream,
                       view_of(h->mesh), n, x0, y0, z0, x, y, z, q, nsteps_arr, nsteps, fe);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

int shp_compare_accel(shp_handle h, int n, const int* status, const double* ax, const double* ay,
                      const double* az, const double* rx, const double* ry, const double* rz, double* out) {
    SHP_REQUIRE(h && out, "shp_compare_accel: null argument");
    SHP_REQUIRE(n >= 0, "shp_compare_accel: n %d < 0", n);
    SHP_REQUIRE(n == 0 || (status && ax && ay && az && rx && ry && rz), "shp_compare_accel: null array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    hipStream_t st = h->stream;
    const int nblk = std::max(1, std::min(kCmpBlocks, (n + 255) / 256));
    // partials: sums and maxima (doubles), then indices and counts (ints), then the result
    SHP_WS(part, double, h, WS_CMP, 3 * (size_t)nblk + 4);
    double* psum = part;
    double* pmax = part + nblk;
    int* pidx = (int*)(part + 2 * nblk);
    int* pcnt = pidx + nblk;
    double* dout = part + 3 * nblk;
    hipLaunchKernelGGL(k_cmp_partial, dim3(nblk), dim3(256), 0, st, n, status, ax, ay, az, rx, ry, rz, psum,
                       pmax, pidx, pcnt);
    hipLaunchKernelGGL(k_cmp_final, dim3(1), dim3(64), 0, st, nblk, psum, pmax, pidx, pcnt, dout);
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipMemcpyAsync(out, dout, 4 * sizeof(double), hipMemcpyDeviceToHost, st));
    SHP_HIP_CHECK(hipStreamSynchronize(st));
    return SHP_OK;
}



// ------------------------------------------------------------ substeps ----

namespace {
struct SubstepRun {
    shp_handle h;
    shp_state* s;
    const shp_step_opts* o;
    const shp_step_opts* sub;
    UpdateParams P, Psub;
    MeshView M;
    SubstepTable T;
    int* level;
    int* eff;
    double *ax, *ay, *az;
    double tpos, tvel;
    dim3 grid, block;
};

int ss_drift(SubstepRun& R) {
    const double hh = R.T.dtl[R.T.M] / 2;
    shp_state* s = R.s;
    R.P.seq = R.h->drift_seq++;
    hipLaunchKernelGGL(k_drift, R.grid, R.block, 0, R.h->stream, R.P, s->n, s->x, s->y, s->z, s->vx, s->vy, s->vz,
                       s->status, s->species, s->t_inj, R.tpos, hh, R.eff);
    SHP_HIP_CHECK(hipGetLastError());
    R.tpos += hh;
    return SHP_OK;
}

int ss_kick(SubstepRun& R, int top) {
    shp_state* s = R.s;
    const int n = s->n;
    hipLaunchKernelGGL(k_ss_status, R.grid, R.block, 0, R.h->stream, n, s->t_inj, R.tpos, R.level, top, s->status,
                       R.eff);
    const shp_step_opts* o = top == 0 ? R.o : R.sub;
    shp_particles p{n, s->x, s->y, s->z, s->q, R.eff, s->species};
    const int rc = o->method == SHP_METHOD_FMM
                       ? shp_fmm_sym_dev(R.h, &p, &o->sym, o->fmm_order, o->leaf_level, o->leaf_size,
                                         o->fmm_ring > 0 ? o->fmm_ring : 2, R.ax, R.ay, R.az)
                   : o->method == SHP_METHOD_BH ? shp_bh_dev(R.h, &p, (float)o->theta, &o->sym, R.ax, R.ay, R.az)
                                                : shp_direct_dev(R.h, &p, o->precision, &o->sym, R.ax, R.ay, R.az);
    if (rc) return rc;
    hipLaunchKernelGGL(k_ss_kick, R.grid, R.block, 0, R.h->stream, R.P, R.M, R.T, n, s->x, s->y, s->z, s->vx, s->vy,
                       s->vz, s->q, s->species, s->t_inj, R.eff, R.level, R.ax, R.ay, R.az, R.tvel);
    SHP_HIP_CHECK(hipGetLastError());
    R.tvel += R.T.dtl[R.T.M];
    return SHP_OK;
}

int ss_K(SubstepRun& R, int L, int top);

int ss_D(SubstepRun& R, int L, bool toggle) {
    if (L == R.T.M) return ss_drift(R);
    int rc = ss_D(R, L + 1, false);
    if (!rc) rc = ss_K(R, L + 1, L + 1);
    if (!rc) rc = ss_D(R, L + 1, true);
    if (!rc && toggle) {
        shp_state* s = R.s;
        hipLaunchKernelGGL(k_ss_crossover, R.grid, R.block, 0, R.h->stream, R.T, L, s->n, s->x, s->y, s->z,
                           s->t_inj, R.tpos, R.level);
        SHP_HIP_CHECK(hipGetLastError());
    }
    return rc;
}

int ss_K(SubstepRun& R, int L, int top) {
    if (L == R.T.M) return ss_kick(R, top);
    int rc = ss_D(R, L + 1, false);
    if (!rc) rc = ss_K(R, L + 1, top);
    if (!rc) rc = ss_D(R, L + 1, true);
    return rc;
}
}  // namespace

int shp_leapfrog_substeps_run(shp_handle h, shp_state* s, const shp_step_opts* o, const shp_step_opts* sub,
                              const shp_substep_level* levels, int max_level, int* level, double t0, double dt,
                              int nsteps, double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && s && o && levels && level && ax && ay && az, "shp_leapfrog_substeps_run: null argument");
    SHP_REQUIRE(max_level >= 0 && max_level < kMaxSubLevels, "shp_leapfrog_substeps_run: max_level %d (0..%d)",
                max_level, kMaxSubLevels - 1);
    SHP_REQUIRE(s->n >= 0 && nsteps >= 0 && dt >= 0, "shp_leapfrog_substeps_run: n=%d steps=%d dt=%g", s->n,
                nsteps, dt);
    SHP_REQUIRE(s->x && s->y && s->z && s->vx && s->vy && s->vz && s->q && s->status && s->species && s->t_inj,
                "shp_leapfrog_substeps_run: null state array");
    if (!sub) sub = o;
    for (const shp_step_opts* q : {o, sub})
        SHP_REQUIRE(q->method == SHP_METHOD_DIRECT || q->method == SHP_METHOD_FMM || q->method == SHP_METHOD_BH,
                    "shp_leapfrog_substeps_run: method %d not available", q->method);
    SHP_REQUIRE(!o->use_efield || h->mesh.ready, "shp_leapfrog_substeps_run: E-field requested but no mesh");
    for (int l = 0; l <= max_level; l++)
        SHP_REQUIRE(levels[l].bc_type >= 0 && levels[l].bc_type <= 3, "shp_leapfrog_substeps_run: level %d bc_type %d",
                    l, levels[l].bc_type);
    if (s->n == 0 || nsteps == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SubstepRun R;
    R.h = h;
    R.s = s;
    R.o = o;
    R.sub = sub;
    R.P = make_params(o);
    attach_outputs(h, R.P);
    R.M = h->mesh.ready ? view_of(h->mesh) : MeshView{};
    R.T = SubstepTable{};
    R.T.M = max_level;
    for (int l = 0; l <= max_level; l++) {
        R.T.type[l] = levels[l].bc_type;
        for (int c = 0; c < 4; c++) R.T.bc[l][c] = levels[l].bc_space[c];
        R.T.bt[l] = levels[l].bc_time;
        R.T.dtl[l] = l == 0 ? dt : R.T.dtl[l - 1] / 3.0;
    }
    R.level = level;
    SHP_WS(eff, int, h, WS_EFF_STATUS, s->n);
    R.eff = eff;
    R.ax = ax;
    R.ay = ay;
    R.az = az;
    R.grid = dim3(cdiv_l(s->n, 256));
    R.block = dim3(256);
    for (int k = 0; k < nsteps; k++) {
        R.tpos = R.tvel = t0 + k * dt;
        int rc = ss_D(R, 0, false);
        if (!rc) rc = ss_K(R, 0, 0);
        if (!rc) rc = ss_D(R, 0, true);
        if (rc) return rc;
    }
    return SHP_OK;
}

// ---------------------------------------------------------- resident state ----

int shp_state_upload_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4,
                          const float* t_inj, int n) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4 && t_inj, "shp_state_upload_host: null argument");
    SHP_REQUIRE(n >= 0, "shp_state_upload_host: n=%d", n);
    SHP_HIP_CHECK(hipSetDevice(h->device));
    if (n > h->res_cap) {
        SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
        if (h->res_f64) SHP_HIP_CHECK(hipFree(h->res_f64));
        if (h->res_i32) SHP_HIP_CHECK(hipFree(h->res_i32));
        if (h->res_aos) SHP_HIP_CHECK(hipFree(h->res_aos));
        if (h->res_e2) SHP_HIP_CHECK(hipFree(h->res_e2));
        if (h->res_ev) SHP_HIP_CHECK(hipFree(h->res_ev));
        if (h->res_evkey) SHP_HIP_CHECK(hipFree(h->res_evkey));
        h->res_f64 = nullptr;
        h->res_i32 = nullptr;
        h->res_aos = nullptr;
        h->res_e2 = nullptr;
        h->res_ev = nullptr;
        h->res_evkey = nullptr;
        h->res_cap = 0;
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_f64, sizeof(double) * 11 * (size_t)n));
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_i32, sizeof(int) * 2 * (size_t)n));
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_aos, sizeof(float4) * 3 * (size_t)n + sizeof(float) * (size_t)n));
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_e2, sizeof(float) * 3 * (size_t)n));
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_ev, sizeof(unsigned int) * kEventWords * (size_t)n));
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_evkey, sizeof(unsigned long long) * (size_t)n));
        h->res_cap = n;
    }
    if (!h->res_evn) SHP_HIP_CHECK(hipMalloc((void**)&h->res_evn, sizeof(int)));
    h->res_n = n;
    h->drift_seq = 0;
    // no kick yet (Efield2 starts at zero, particles.cpp:1040-1041) and no event
    hipLaunchKernelGGL(k_ss_fill, dim3(1), dim3(64), 0, h->stream, h->res_evn, 1, 0);
    if (n == 0) return SHP_OK;
    hipLaunchKernelGGL(k_ss_fill, dim3(cdiv_l(3 * (long)n, 256)), dim3(256), 0, h->stream, (int*)h->res_e2, 3 * n, 0);
    float4* dpos = h->res_aos;
    float4* dvel = dpos + n;
    float4* dacc = dvel + n;
    float* dti = (float*)(dacc + n);
    SHP_HIP_CHECK(hipMemcpyAsync(dpos, pos4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dvel, vel4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dacc, acc4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dti, t_inj, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
    double* f = h->res_f64;
    hipLaunchKernelGGL(k_res_from_aos, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, dpos, dvel, dacc, dti, n,
                       f, f + n, f + 2 * (size_t)n, f + 3 * (size_t)n, f + 4 * (size_t)n, f + 5 * (size_t)n,
                       f + 6 * (size_t)n, f + 7 * (size_t)n, h->res_i32, h->res_i32 + n, f + 8 * (size_t)n);
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));  // the host arrays may change after return
    return SHP_OK;
}

// Marks the leapfrog launches of a resident run, which record each kick's
// field and the boundary-condition events.
struct ResidentOutputs {
    shp_handle h;
    explicit ResidentOutputs(shp_handle hh) : h(hh) { h->res_events_on = true; }
    ~ResidentOutputs() { h->res_events_on = false; }
};

int shp_state_run(shp_handle h, const shp_step_opts* o, double t0, double dt, int nsteps) {
    SHP_REQUIRE(h && o, "shp_state_run: null argument");
    SHP_REQUIRE(h->res_f64 || h->res_n == 0, "shp_state_run: no resident state (shp_state_upload_host)");
    const int n = h->res_n;
    if (n == 0 || nsteps == 0) return SHP_OK;
    double* f = h->res_f64;
    shp_state s{n, f, f + n, f + 2 * (size_t)n, f + 3 * (size_t)n, f + 4 * (size_t)n, f + 5 * (size_t)n,
                f + 6 * (size_t)n, h->res_i32, h->res_i32 + n, f + 7 * (size_t)n};
    ResidentOutputs on(h);
    return shp_leapfrog_run(h, &s, o, t0, dt, nsteps, f + 8 * (size_t)n, f + 9 * (size_t)n, f + 10 * (size_t)n);
}

int shp_state_run_substeps(shp_handle h, const shp_step_opts* o, const shp_step_opts* sub,
                           const shp_substep_level* levels, int max_level, int reset_levels, double t0, double dt,
                           int nsteps) {
    SHP_REQUIRE(h && o && levels, "shp_state_run_substeps: null argument");
    SHP_REQUIRE(h->res_f64 || h->res_n == 0, "shp_state_run_substeps: no resident state (shp_state_upload_host)");
    SHP_REQUIRE(max_level >= 0 && max_level < kMaxSubLevels, "shp_state_run_substeps: max_level %d", max_level);
    const int n = h->res_n;
    if (n == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    if (n > h->res_level_cap) {
        SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
        if (h->res_level) SHP_HIP_CHECK(hipFree(h->res_level));
        h->res_level = nullptr;
        h->res_level_cap = 0;
        SHP_HIP_CHECK(hipMalloc((void**)&h->res_level, sizeof(int) * (size_t)n));
        h->res_level_cap = n;
        reset_levels = 1;
    }
    if (reset_levels)  // initializeParticleLevels: every ion at the finest level
        hipLaunchKernelGGL(k_ss_fill, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, h->res_level, n, max_level);
    double* f = h->res_f64;
    shp_state s{n, f, f + n, f + 2 * (size_t)n, f + 3 * (size_t)n, f + 4 * (size_t)n, f + 5 * (size_t)n,
                f + 6 * (size_t)n, h->res_i32, h->res_i32 + n, f + 7 * (size_t)n};
    ResidentOutputs on(h);
    return shp_leapfrog_substeps_run(h, &s, o, sub, levels, max_level, h->res_level, t0, dt, nsteps,
                                     f + 8 * (size_t)n, f + 9 * (size_t)n, f + 10 * (size_t)n);
}

int shp_state_download_host(shp_handle h, float* pos4, float* vel4, float* acc4) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4, "shp_state_download_host: null argument");
    const int n = h->res_n;
    if (n == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    float4* dpos = h->res_aos;
    float4* dvel = dpos + n;
    float4* dacc = dvel + n;
    const double* f = h->res_f64;
    hipLaunchKernelGGL(k_res_to_aos, dim3(cdiv_l(n, 256)), dim3(256), 0, h->stream, f, f + n, f + 2 * (size_t)n,
                       f + 3 * (size_t)n, f + 4 * (size_t)n, f + 5 * (size_t)n, f + 8 * (size_t)n,
                       f + 9 * (size_t)n, f + 10 * (size_t)n, h->res_i32, n, dpos, dvel, dacc);
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipMemcpyAsync(pos4, dpos, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(vel4, dvel, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(acc4, dacc, sizeof(float4) * n, hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

}  // extern "C"
