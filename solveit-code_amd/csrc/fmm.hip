// Fast multipole method on gfx950 for the softened Coulomb kernel
// G(r) = (|r|^2 + eps^2)^(-1/2), FP32 expansion and pair arithmetic on FP64
// state.
//
// Replaces MultipoleMethodForceCalculation::updateAcceleration and its passes
// (multipole_method_force_calculation.cpp:148-522, fmm_cpukernel.cpp:103-1317).
// The reference translates spherical-harmonic expansions of the plain 1/r
// kernel and softens only the centre distance in M2L (fmm_cpukernel.cpp:832);
// here the softened kernel itself is expanded in Cartesian Taylor series, so
// the far field converges to the softened direct sum.
//
// Tree: cube root around the FP64 bounding box; 63-bit Morton keys (21 bits per
// axis, x -> bit 3i, y -> 3i+1, z -> 3i+2, multipole_method_functions.cpp:10-26
// uses another interleave) sorted once; every box at every level is a
// contiguous run of the sorted particles.  Uniform leaf level L.  Near field:
// leaf pairs whose cell coordinates differ by at most `ring` per axis.  Far
// field at level l: children of the parent's near boxes that are not near.
//
// Expansions (lengths in units of the box half size h of the level):
//   P2M  M_a = sum q v^a                      v = (y - c)/h
//   M2M  M'_a += sum_{k<=a} C(a,k) 2^-|a| s(o, a-k) M_k     (s: octant sign)
//   M2L  L_b += sum_{|a|+|b|<=p} C(a+b,a) (-1)^|a| T_{a+b}(2 D, eps/h) M_a
//   L2L  L'_k += 1/2 sum_{b>=k} C(b,k) 2^-|b| s(o, b-k) L_b
//   L2P  a = -K q/m (1/h^2) grad_u sum_b L_b u^b       u = (x - c)/h
// T_g are the Taylor coefficients of G at the integer box offset D (R = 2D in
// units of h), tabulated per level by the recurrence
//   |g| s T_g = -(2|g|-1) sum_i R_i T_{g-e_i} - (|g|-1) sum_i T_{g-2e_i},
//   s = |R|^2 + (eps/h)^2.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>
#include <vector>

#include "shp_internal.h"

namespace shp {

static inline int fdiv_up(long a, long b) { return (int)((a + b - 1) / b); }

constexpr int kMaxP = 12;
constexpr int kMaxNc = (kMaxP + 1) * (kMaxP + 2) * (kMaxP + 3) / 6;  // 455
constexpr int kMaxLevel = 21;
constexpr int kFmmThreads = 256;
constexpr int kM2LBatch = 4;
constexpr int kShiftLanes = 64;  // one wave per parent box in M2M / L2L
constexpr int kShiftUnroll = 16;  // shift terms loaded per batch (the lists are padded to a multiple)

// Index of the coefficient (a, b, c) in the degree-major order (within a
// degree: a descending, then b descending) and first index of degree n.
constexpr int cidx(int a, int b, int c) {
    return (a + b + c) * (a + b + c + 1) * (a + b + c + 2) / 6 + (b + c) * (b + c + 1) / 2 + c;
}
constexpr int dstart(int n) { return n * (n + 1) * (n + 2) / 6; }

// ------------------------------------------------------------- host tables ----

struct FmmHostTables {
    int p = -1, nc = 0;
    std::vector<int4> mi;  // a, b, c, degree
    int idx[kMaxP + 1][kMaxP + 1][kMaxP + 1];
    // M2L tasks: pair ranges, one accumulator each
    std::vector<int> pair_alpha, pair_gamma;
    std::vector<float> pair_w;
    std::vector<int2> task;        // [pair begin, pair end)
    std::vector<int> task_beta;
    std::vector<int2> beta_tasks;  // [task begin, task end) per beta
    // M2M / L2L term lists of the 64 lanes of a shift workgroup, stored
    // lane-interleaved (entry t of lane l at t * 64 + l) and padded with zero
    // terms to the longest lane: {src | dst << 12 | last << 24, coefficient}.
    // A lane accumulates coef * S[src] and stores to dst when `last` is set.
    // M2M: src = class * nc + k in the octant transform, dst = output a;
    // L2L: src = parent coefficient b, dst = class * nc + k.
    std::vector<int2> m2m_term, l2l_term;
    int m2m_nt = 0, l2l_nt = 0;
};

static double binom_d(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

static void build_tables(FmmHostTables& t, int p) {
    t.p = p;
    t.mi.clear();
    for (int n = 0; n <= p; n++)
        for (int a = n; a >= 0; a--)
            for (int b = n - a; b >= 0; b--) {
                int c = n - a - b;
                t.idx[a][b][c] = (int)t.mi.size();
                t.mi.push_back(make_int4(a, b, c, n));
            }
    t.nc = (int)t.mi.size();
    const int nc = t.nc;
    // M2L pairs grouped by beta, split into tasks of at most G pairs so that
    // the number of tasks fits one workgroup.
    std::vector<std::vector<int>> pa(nc), pg(nc);
    std::vector<std::vector<float>> pw(nc);
    long npairs = 0;
    for (int be = 0; be < nc; be++) {
        int4 B = t.mi[be];
        for (int al = 0; al < nc; al++) {
            int4 A = t.mi[al];
            if (A.w + B.w > p) continue;
            double w = binom_d(A.x + B.x, A.x) * binom_d(A.y + B.y, A.y) * binom_d(A.z + B.z, A.z);
            if (A.w & 1) w = -w;
            pa[be].push_back(al);
            pg[be].push_back(t.idx[A.x + B.x][A.y + B.y][A.z + B.z]);
            pw[be].push_back((float)w);
            npairs++;
        }
    }
    int G = 1;
    for (;; G++) {
        long ntask = 0;
        for (int be = 0; be < nc; be++) ntask += (pa[be].size() + G - 1) / G;
        if (ntask <= 2 * kFmmThreads) break;
    }
    t.pair_alpha.clear();
    t.pair_gamma.clear();
    t.pair_w.clear();
    t.task.clear();
    t.task_beta.clear();
    t.beta_tasks.assign(nc, make_int2(0, 0));
    for (int be = 0; be < nc; be++) {
        int tb = (int)t.task.size();
        int n = (int)pa[be].size();
        for (int s = 0; s < n; s += G) {
            int e = std::min(n, s + G);
            int pb = (int)t.pair_alpha.size();
            for (int k = s; k < e; k++) {
                t.pair_alpha.push_back(pa[be][k]);
                t.pair_gamma.push_back(pg[be][k]);
                t.pair_w.push_back(pw[be][k]);
            }
            t.task.push_back(make_int2(pb, (int)t.pair_alpha.size()));
            t.task_beta.push_back(be);
        }
        t.beta_tasks[be] = make_int2(tb, (int)t.task.size());
    }
    (void)npairs;
    // Shift term lists: outputs (M2M: a; L2L: (k, class of b)) are dealt
    // longest first onto the least loaded lane; each output is a run of terms
    // ending with `last`.
    struct Term { int src, dst; float c; };
    auto deal = [](std::vector<std::vector<Term>>& outs, std::vector<int2>& table, int& nt) {
        std::vector<int> order(outs.size());
        for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return outs[x].size() > outs[y].size(); });
        std::vector<std::vector<int2>> lane(kShiftLanes);
        std::vector<size_t> load(kShiftLanes, 0);
        for (int i : order) {
            if (outs[i].empty()) continue;
            const int ln = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            for (size_t j = 0; j < outs[i].size(); j++) {
                const Term& e = outs[i][j];
                const int last = j + 1 == outs[i].size() ? 1 : 0;
                int bits;
                std::memcpy(&bits, &e.c, sizeof(bits));
                lane[ln].push_back(make_int2(e.src | (e.dst << 12) | (last << 24), bits));
            }
            load[ln] += outs[i].size();
        }
        // a whole number of kShiftUnroll-term batches per lane; the padding
        // terms have coefficient 0 and no `last` flag
        nt = (int)*std::max_element(load.begin(), load.end());
        nt = (nt + kShiftUnroll - 1) / kShiftUnroll * kShiftUnroll;
        table.assign((size_t)nt * kShiftLanes, make_int2(0, 0));
        for (int ln = 0; ln < kShiftLanes; ln++)
            for (size_t j = 0; j < lane[ln].size(); j++) table[j * kShiftLanes + ln] = lane[ln][j];
    };
    {
        // M2M: M'_a = 2^-|a| sum_{k<=a} C(a,k) W_k[class(a-k)]
        std::vector<std::vector<Term>> outs(nc);
        for (int a = 0; a < nc; a++) {
            const int4 A = t.mi[a];
            for (int kx = 0; kx <= A.x; kx++)
                for (int ky = 0; ky <= A.y; ky++)
                    for (int kz = 0; kz <= A.z; kz++) {
                        const int cls = ((A.x - kx) & 1) | (((A.y - ky) & 1) << 1) | (((A.z - kz) & 1) << 2);
                        const double c = binom_d(A.x, kx) * binom_d(A.y, ky) * binom_d(A.z, kz) *
                                         std::ldexp(1.0, -A.w);
                        outs[a].push_back({cls * nc + t.idx[kx][ky][kz], a, (float)c});
                    }
        }
        deal(outs, t.m2m_term, t.m2m_nt);
    }
    {
        // L2L: Q_r[k] = 1/2 sum_{b>=k, class(b)=r} C(b,k) 2^-|b| L_b
        std::vector<std::vector<Term>> outs(8 * (size_t)nc);
        for (int k = 0; k < nc; k++) {
            const int4 K = t.mi[k];
            for (int b = 0; b < nc; b++) {
                const int4 B = t.mi[b];
                if (B.x < K.x || B.y < K.y || B.z < K.z) continue;
                const int cls = (B.x & 1) | ((B.y & 1) << 1) | ((B.z & 1) << 2);
                const double c = 0.5 * binom_d(B.x, K.x) * binom_d(B.y, K.y) * binom_d(B.z, K.z) *
                                 std::ldexp(1.0, -B.w);
                outs[(size_t)cls * nc + k].push_back({b, cls * nc + k, (float)c});
            }
        }
        deal(outs, t.l2l_term, t.l2l_nt);
    }
}

struct FmmDevTables {
    int p = -1, nc = 0, ntask = 0;
    long npairs = 0;  // (alpha, beta) pairs of one M2L
    int4* mi = nullptr;
    int* pair_alpha = nullptr;
    int* pair_gamma = nullptr;
    float* pair_w = nullptr;
    int2* task = nullptr;
    int2* beta_tasks = nullptr;
    int2* m2m_term = nullptr;
    int2* l2l_term = nullptr;
    int m2m_nt = 0, l2l_nt = 0;
    float* mscale = nullptr;
    float* dscale = nullptr;
    float* lscale = nullptr;
    float* zeros = nullptr;
    // far-offset list of the tiled M2L (ring 2) and its chunk bounds
    short* olist = nullptr;
    int olist_n = 0, olist_s4 = 0, olist_s5 = 0;
    int* cbound = nullptr;
    int cbound_key = -1;  // (shell orders, chunks) the bounds were made for
};

static std::map<shp_handle, FmmDevTables> g_tables;

template <class T>
static int upload(T** dst, const std::vector<T>& v) {
    if (*dst) (void)hipFree(*dst);
    *dst = nullptr;
    SHP_HIP_CHECK(hipMalloc((void**)dst, sizeof(T) * std::max<size_t>(1, v.size())));
    if (!v.empty()) SHP_HIP_CHECK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return SHP_OK;
}

static int get_tables(shp_handle h, int p, FmmDevTables** out) {
    FmmDevTables& d = g_tables[h];
    if (d.p != p) {
        FmmHostTables t;
        build_tables(t, p);
        int rc = 0;
        rc |= upload(&d.mi, t.mi);
        rc |= upload(&d.pair_alpha, t.pair_alpha);
        rc |= upload(&d.pair_gamma, t.pair_gamma);
        rc |= upload(&d.pair_w, t.pair_w);
        rc |= upload(&d.task, t.task);
        rc |= upload(&d.beta_tasks, t.beta_tasks);
        rc |= upload(&d.m2m_term, t.m2m_term);
        rc |= upload(&d.l2l_term, t.l2l_term);
        d.m2m_nt = t.m2m_nt;
        d.l2l_nt = t.l2l_nt;
        // scalings of the unrolled M2L: Mhat_a = (-1)^|a| M_a / a!,
        // D_g = g! T_g, L_b += Lhat_b / b!
        std::vector<float> ms(t.nc), ds(t.nc), ls(t.nc), zero(t.nc, 0.0f);
        auto fact = [](int k) { double f = 1; for (int i = 2; i <= k; i++) f *= i; return f; };
        for (int k = 0; k < t.nc; k++) {
            const int4 m = t.mi[k];
            const double f = fact(m.x) * fact(m.y) * fact(m.z);
            ms[k] = (float)((m.w & 1 ? -1.0 : 1.0) / f);
            ds[k] = (float)f;
            ls[k] = (float)(1.0 / f);
        }
        rc |= upload(&d.mscale, ms);
        rc |= upload(&d.dscale, ds);
        rc |= upload(&d.lscale, ls);
        rc |= upload(&d.zeros, zero);
        if (rc) return SHP_ERR_HIP;
        d.p = p;
        d.nc = t.nc;
        d.ntask = (int)t.task.size();
        d.npairs = (long)t.pair_alpha.size();
    }
    *out = &d;
    return SHP_OK;
}

static std::map<shp_handle, int> g_last_level;
static int& last_leaf_level(shp_handle h) { return g_last_level[h]; }

void fmm_release(shp_handle h) {
    g_last_level.erase(h);
    auto it = g_tables.find(h);
    if (it == g_tables.end()) return;
    FmmDevTables& d = it->second;
    void* ptrs[] = {d.mi, d.pair_alpha, d.pair_gamma, d.pair_w, d.task, d.beta_tasks, d.m2m_term,
                    d.l2l_term,
                    d.mscale, d.dscale, d.lscale, d.zeros, d.olist, d.cbound};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    g_tables.erase(it);
}

// ------------------------------------------------------------------ keys ----

__device__ __forceinline__ uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
__device__ __forceinline__ uint32_t compact21(uint64_t v) {
    v &= 0x1249249249249249ULL;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ULL;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00fULL;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffULL;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
    v = (v ^ (v >> 32)) & 0x1fffff;
    return (uint32_t)v;
}

__global__ void k_fmm_bbox_partial(const double* __restrict__ X, const double* __restrict__ Y,
                                   const double* __restrict__ Z, const int* __restrict__ status,
                                   int n, double* __restrict__ part) {
    __shared__ double red[6][kFmmThreads];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int s = status[i];
        if (!(s == 1 || s == 2)) continue;
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
    for (int s = kFmmThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + s]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Root cube (x0, y0, z0, side): side = largest extent times (1 + 2^-20),
// centred on the bounding box; root[4..9] = bounding box min, max.  One wave.
__global__ __launch_bounds__(64) void k_fmm_root(const double* __restrict__ part, int nparts,
                                                 double* __restrict__ root) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int b = threadIdx.x; b < nparts; b += 64)
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], part[b * 6 + c]);
            mx[c] = fmax(mx[c], part[b * 6 + 3 + c]);
        }
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], __shfl_xor(mn[c], off));
            mx[c] = fmax(mx[c], __shfl_xor(mx[c], off));
        }
    if (threadIdx.x != 0) return;
    double side = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    if (!(side > 0)) side = 1e-9;
    side = side * (1.0 + 0x1p-20);
    for (int c = 0; c < 3; c++) root[c] = 0.5 * (mn[c] + mx[c]) - 0.5 * side;
    root[3] = side;
    for (int c = 0; c < 3; c++) {
        root[4 + c] = mn[c];
        root[7 + c] = mx[c];
    }
}

// Level-21 Morton key of each active particle; inactive ones sort last.
__global__ void k_fmm_keys(const double* __restrict__ X, const double* __restrict__ Y,
                           const double* __restrict__ Z, const int* __restrict__ status, int n,
                           const double* __restrict__ root, uint64_t* __restrict__ keys,
                           int* __restrict__ vals) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = status[i];
    uint64_t k = ~0ULL;
    if (s == 1 || s == 2) {
        const double sc = 2097152.0 / root[3];
        const double v[3] = {X[i], Y[i], Z[i]};
        k = 0;
        for (int c = 0; c < 3; c++) {
            const double f = floor((v[c] - root[c]) * sc);
            int64_t iv = (int64_t)f;
            if (f < 0) iv = 0;
            if (iv > 2097151) iv = 2097151;
            k |= spread21((uint64_t)iv) << c;
        }
    }
    keys[i] = k;
    vals[i] = i;
}

// Sorted particle records relative to nothing yet: double position, charge.
__global__ void k_fmm_gather(const int* __restrict__ perm, int na, const double* __restrict__ X,
                             const double* __restrict__ Y, const double* __restrict__ Z,
                             const double* __restrict__ Q, double4* __restrict__ sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    const int j = perm[i];
    sp[i] = make_double4(X[j], Y[j], Z[j], Q[j]);
}

// 1 for a sorted ion that is a target (status 1), for the prefix sums of
// box_has_target; entry na is 0.
__global__ void k_fmm_tflag(const int* __restrict__ perm, int na, const int* __restrict__ status,
                            int* __restrict__ tf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > na) return;
    tf[i] = i < na && status[perm[i]] == 1 ? 1 : 0;
}

// Number of active particles (keys != ~0 sort last) and, per level, the number
// of particles that open a new box there: bin[lmin] counts particles whose key
// first differs from the previous one at level lmin (so they open boxes at all
// levels >= lmin).  counts = {na, bin[0..21]}.
__global__ __launch_bounds__(256) void k_fmm_level_hist(const uint64_t* __restrict__ keys, int n,
                                                        int* __restrict__ counts) {
    __shared__ int bins[kMaxLevel + 1];
    if (threadIdx.x <= kMaxLevel) bins[threadIdx.x] = 0;
    __syncthreads();
    // grid-stride over the keys with a bounded grid: the per-block bins meet
    // in a few hundred global atomics (same-address atomics serialise in L2)
    const int n256 = (n + 255) & ~255;  // whole waves take every trip
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n256; i += gridDim.x * blockDim.x) {
        int lmin = kMaxLevel + 1;  // nothing to count
        if (i < n) {
            const uint64_t k = keys[i];
            if (k != ~0ULL) {
                if (i + 1 == n || keys[i + 1] == ~0ULL) counts[0] = i + 1;
                lmin = 0;
                if (i > 0) {
                    const uint64_t x = k ^ keys[i - 1];
                    lmin = x == 0 ? kMaxLevel + 1 : kMaxLevel - (63 - __clzll((long long)x)) / 3;
                }
            }
        }
        // one LDS atomic per distinct level in the wave (most keys of a wave
        // share one), instead of one per key on the same few bins
        unsigned long long rem = __ballot(lmin <= kMaxLevel);
        while (rem) {
            const int src = __ffsll((long long)rem) - 1;
            const int lb = __shfl(lmin, src);
            const unsigned long long m = __ballot(lmin == lb);
            if ((int)(threadIdx.x & 63) == src) atomicAdd(&bins[lb], __popcll(m));
            rem &= ~m;
        }
    }
    __syncthreads();
    if (threadIdx.x <= kMaxLevel && bins[threadIdx.x]) atomicAdd(&counts[1 + threadIdx.x], bins[threadIdx.x]);
}

__global__ void k_fmm_u64fill(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

__global__ void k_fmm_ifill(int* __restrict__ p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Box starts at level l: flags of prefix changes.
__global__ void k_fmm_flags(const uint64_t* __restrict__ keys, int na, int sh, char* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    f[i] = (i == 0 || (keys[i - 1] >> sh) != (keys[i] >> sh)) ? 1 : 0;
}

// Per box of the leaf-most level built: key at its level and particle range end.
__global__ void k_fmm_box_info(const uint64_t* __restrict__ keys, const int* __restrict__ start,
                               int nb, int na, int sh, uint64_t* __restrict__ bkey,
                               int* __restrict__ bend) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int s = start[b];
    bkey[b] = keys[s] >> sh;
    bend[b] = b + 1 < nb ? start[b + 1] : na;
}

// Level-l boxes from the level-(l+1) boxes: a child box starts a parent box
// when its parent key differs from the previous child's.
__global__ void k_fmm_parent_flags(const uint64_t* __restrict__ ckey, int ncb, char* __restrict__ f) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= ncb) return;
    f[b] = (b == 0 || (ckey[b - 1] >> 3) != (ckey[b] >> 3)) ? 1 : 0;
}

// Per parent box k with first child cfirst[k]: key, particle start and end.
__global__ void k_fmm_parent_info(const int* __restrict__ cfirst, int nb, int na, const uint64_t* __restrict__ ckey,
                                  const int* __restrict__ cstart, uint64_t* __restrict__ bkey,
                                  int* __restrict__ start, int* __restrict__ bend) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nb) return;
    const int c = cfirst[k];
    bkey[k] = ckey[c] >> 3;
    start[k] = cstart[c];
    bend[k] = k + 1 < nb ? cstart[cfirst[k + 1]] : na;
}

// The levels lhi, lhi - 1, ..., 0 built one after another from the level
// below by a single workgroup (these levels hold few boxes; one launch
// instead of a device-wide select per level).
constexpr int kUpperMaxChildren = 16384;  // children of the top level built this way, at most
struct UpperLevels {
    int lhi, na;
    uint64_t* key[kMaxLevel + 2];
    int* start[kMaxLevel + 2];
    int* end[kMaxLevel + 2];
    int* cfirst[kMaxLevel + 2];
    int nb[kMaxLevel + 2];
};

__global__ __launch_bounds__(1024) void k_fmm_upper_levels(const UpperLevels U) {
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int l = U.lhi; l >= 0; l--) {
        const int ncb = U.nb[l + 1];
        const uint64_t* ck = U.key[l + 1];
        const int* cs = U.start[l + 1];
        int carry = 0;  // parent boxes found so far (uniform)
        for (int b0 = 0; b0 < ncb; b0 += 1024) {
            const int b = b0 + (int)threadIdx.x;
            const bool f = b < ncb && (b == 0 || (ck[b - 1] >> 3) != (ck[b] >> 3));
            const unsigned long long m = __ballot(f);
            if (lane == 0) wsum[w] = __popcll(m);
            __syncthreads();
            int before = 0, tot = 0;
            for (int k = 0; k < 16; k++) {
                const int v = wsum[k];
                before += k < w ? v : 0;
                tot += v;
            }
            if (f) {
                const int k = carry + before + __popcll(m & ((1ull << lane) - 1ull));
                U.cfirst[l][k] = b;
                U.key[l][k] = ck[b] >> 3;
                U.start[l][k] = cs[b];
                if (k > 0) U.end[l][k - 1] = cs[b];
            }
            carry += tot;
            __syncthreads();  // wsum is reused
        }
        if (threadIdx.x == 0 && carry > 0) U.end[l][carry - 1] = U.na;
        __syncthreads();  // level l is complete before level l - 1 reads it
    }
}

__device__ __forceinline__ int find_key(const uint64_t* __restrict__ bkey, int nb, uint64_t k) {
    int lo = 0, hi = nb - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        const uint64_t v = bkey[mid];
        if (v == k) return mid;
        if (v < k) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

struct Level {
    int nb;
    const uint64_t* key;
    const int* start;
    const int* end;
    const int* cfirst;  // nullptr at the leaf level
    float* M;
    float* Lc;
    int ld;  // row stride of M and Lc (coefficients padded to whole 128-byte lines)
    // prefix sums over the sorted ions of "is a target" (status 1): box b holds
    // a target iff tpre[end[b]] > tpre[start[b]]; nullptr: every box does
    const int* tpre;
};

// Boxes without targets (source-only ions: symmetry images, or the ions of
// other levels in a multi-rate step) still need their multipoles, but no
// local expansion and no near field.
__device__ __forceinline__ bool box_has_target(const Level& lv, int b) {
    return !lv.tpre || lv.tpre[lv.end[b]] > lv.tpre[lv.start[b]];
}

__device__ __forceinline__ void box_center(const double* __restrict__ root, int l, uint64_t key,
                                           double c[3], double& h) {
    const double bs = root[3] / (double)(1 << l);
    for (int d = 0; d < 3; d++) c[d] = root[d] + ((double)compact21(key >> d) + 0.5) * bs;
    h = 0.5 * bs;
}

// ------------------------------------------------------------------- P2M ----

// One workgroup per leaf; powers of v per particle staged in LDS.
__global__ __launch_bounds__(kFmmThreads) void k_fmm_p2m(
    Level lv, int l, const double* __restrict__ root, const double4* __restrict__ sp, int nc,
    int p, const int4* __restrict__ mi) {
    __shared__ float pw[64][3][kMaxP + 1];
    __shared__ float qs[64];
    const int b = blockIdx.x;
    double c[3], h;
    box_center(root, l, lv.key[b], c, h);
    const double ih = 1.0 / h;
    const int s0 = lv.start[b], s1 = lv.end[b];
    float acc0 = 0.0f, acc1 = 0.0f;
    const int a0 = threadIdx.x, a1 = threadIdx.x + kFmmThreads;
    const int4 A0 = a0 < nc ? mi[a0] : make_int4(0, 0, 0, 0);
    const int4 A1 = a1 < nc ? mi[a1] : make_int4(0, 0, 0, 0);
    for (int base = s0; base < s1; base += 64) {
        __syncthreads();
        if (threadIdx.x < 64) {
            const int i = base + threadIdx.x;
            if (i < s1) {
                const double4 r = sp[i];
                const float v[3] = {(float)((r.x - c[0]) * ih), (float)((r.y - c[1]) * ih),
                                    (float)((r.z - c[2]) * ih)};
                for (int d = 0; d < 3; d++) {
                    float t = 1.0f;
                    for (int k = 0; k <= p; k++) {
                        pw[threadIdx.x][d][k] = t;
                        t *= v[d];
                    }
                }
                qs[threadIdx.x] = (float)r.w;
            }
        }
        __syncthreads();
        const int cnt = min(64, s1 - base);
        if (a0 < nc)
            for (int j = 0; j < cnt; j++)
                acc0 = fmaf(qs[j], pw[j][0][A0.x] * pw[j][1][A0.y] * pw[j][2][A0.z], acc0);
        if (a1 < nc)
            for (int j = 0; j < cnt; j++)
                acc1 = fmaf(qs[j], pw[j][0][A1.x] * pw[j][1][A1.y] * pw[j][2][A1.z], acc1);
    }
    if (a0 < nc) lv.M[(size_t)b * lv.ld + a0] = acc0;
    if (a1 < nc) lv.M[(size_t)b * lv.ld + a1] = acc1;
}

// ------------------------------------------------------------------- M2M ----

// The octant sign s(o, g) = prod over axes with child bit 0 of (-1)^g_axis
// depends on g only through its parity class, so the children of a parent are
// first combined by a Walsh-Hadamard transform over the octant index:
//   W_k[r] = sum_children s(o, r) M_k(child)        (r = parity class)
//   M'_a   = 2^-|a| sum_{k<=a} C(a,k) W_k[class(a-k)]
// which makes the shift cost one triangular sum per parent instead of one per
// child.  One wave per parent; the triangular sums run from flat per-lane term
// lists (uniform trip count, no divergent nested loops).
__global__ __launch_bounds__(kShiftLanes) void k_fmm_m2m(Level par, Level ch, int nc,
                                                         const int2* __restrict__ term, int nt) {
    extern __shared__ float sm[];
    float* W = sm;  // [8][nc]: child moments by octant, then by parity class
    const int b = blockIdx.x, t = threadIdx.x;
    const int c0 = par.cfirst[b];
    const int c1 = b + 1 < par.nb ? par.cfirst[b + 1] : ch.nb;
    for (int i = t; i < 8 * nc; i += kShiftLanes) W[i] = 0.0f;
    __syncthreads();
    for (int c = c0; c < c1; c++) {
        const int o = (int)(ch.key[c] & 7);
        const float* Mc = ch.M + (size_t)c * ch.ld;
        for (int k = t; k < nc; k += kShiftLanes) W[o * nc + k] = Mc[k];
    }
    __syncthreads();
    for (int k = t; k < nc; k += kShiftLanes) {
        float u[8];
#pragma unroll
        for (int o = 0; o < 8; o++) u[o] = W[o * nc + k];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int o = 0; o < 8; o++)
                if (!(o & (1 << d))) {
                    const float x = u[o], y = u[o | (1 << d)];
                    u[o] = x + y;            // class bit 0: sum over the axis
                    u[o | (1 << d)] = y - x;  // class bit 1: upper minus lower child
                }
#pragma unroll
        for (int o = 0; o < 8; o++) W[o * nc + k] = u[o];
    }
    __syncthreads();
    float* Mp = par.M + (size_t)b * par.ld;
    float acc = 0.0f;
    // the term loads of a batch are issued together, ahead of its chain of
    // dependent multiply-adds
    for (int j0 = 0; j0 < nt; j0 += kShiftUnroll) {
        int2 e[kShiftUnroll];
#pragma unroll
        for (int u = 0; u < kShiftUnroll; u++) e[u] = term[(j0 + u) * kShiftLanes + t];
#pragma unroll
        for (int u = 0; u < kShiftUnroll; u++) {
            acc = fmaf(__int_as_float(e[u].y), W[e[u].x & 0xfff], acc);
            if (e[u].x >> 24) {
                Mp[(e[u].x >> 12) & 0xfff] = acc;
                acc = 0.0f;
            }
        }
    }
}

// ------------------------------------------------------------- T tables ----

// T[level][offset][nc] for integer offsets D in [-W, W]^3 (R = 2D).  One
// wave per (level, offset): the coefficients of degree n depend only on those
// of degrees n-1 and n-2, so the lanes compute one degree at a time in LDS.
__global__ __launch_bounds__(64) void k_fmm_ttable(int p, int nc, const int4* __restrict__ mi, int W, int L0,
                                                   int L, const double* __restrict__ root, double eps2,
                                                   float* __restrict__ T) {
    __shared__ double Td[kMaxNc];
    const int nside = 2 * W + 1, noff = nside * nside * nside;
    const int t = blockIdx.x;
    if (t >= noff * (L - L0 + 1)) return;  // uniform over the workgroup
    const int l = L0 + t / noff, o = t % noff;
    const int dx = o % nside - W, dy = (o / nside) % nside - W, dz = o / (nside * nside) - W;
    const double h = 0.5 * root[3] / (double)(1 << l);
    const double R[3] = {2.0 * dx, 2.0 * dy, 2.0 * dz};
    const double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps2 / (h * h);
    if (threadIdx.x == 0) Td[0] = 1.0 / sqrt(s);
    __syncthreads();
    // mi is ordered by degree, and within a degree by (a desc, b desc); index of
    // (a,b,c) with n=a+b+c is n(n+1)(n+2)/6 + (n-a)(n-a+1)/2 + (n-a-b).
    for (int n = 1; n <= p; n++) {
        for (int k = dstart(n) + (int)threadIdx.x; k < dstart(n + 1); k += 64) {
            const int4 g = mi[k];
            const int gg[3] = {g.x, g.y, g.z};
            double acc = 0.0;
            for (int i = 0; i < 3; i++) {
                for (int dec = 1; dec <= 2; dec++) {
                    if (gg[i] < dec) continue;
                    int hh[3] = {gg[0], gg[1], gg[2]};
                    hh[i] -= dec;
                    const int m = hh[0] + hh[1] + hh[2];
                    const int id = m * (m + 1) * (m + 2) / 6 + (m - hh[0]) * (m - hh[0] + 1) / 2 +
                                   (m - hh[0] - hh[1]);
                    acc += dec == 1 ? (2 * n - 1) * R[i] * Td[id] : (n - 1) * Td[id];
                }
            }
            Td[k] = -acc / (n * s);
        }
        __syncthreads();
    }
    float* out = T + (size_t)t * nc;
    for (int k = threadIdx.x; k < nc; k += 64) out[k] = (float)Td[k];
}

// ------------------------------------------------------------------- M2L ----

// One workgroup per target box: candidate sources are the children of the
// parent's near boxes that are not near the target, listed in parent order by
// wave 0 with ballot compaction (deterministic).  Up to kM2LBatch sources are
// staged in LDS at a time; each lane runs up to two tasks (slices of the
// (beta, alpha) pairs of one beta) over them, and the task partials of each
// beta are summed in task order.
__global__ __launch_bounds__(kFmmThreads) void k_fmm_m2l(
    Level tl, Level pl, int l, int ring, int nc, int ntask, const int2* __restrict__ task,
    const int* __restrict__ pa, const int* __restrict__ pg, const float* __restrict__ pw,
    const int2* __restrict__ beta_tasks, const float* __restrict__ Ttab, int W,
    unsigned long long* __restrict__ stats) {
    __shared__ int src[1024];
    __shared__ int srcoff[1024];
    __shared__ int nsrc;
    __shared__ float Ms[kM2LBatch][kMaxNc];
    __shared__ float Ts[kM2LBatch][kMaxNc];
    __shared__ float part[2 * kFmmThreads];
    __shared__ int pcnt[128];
    __shared__ int pfirst[128];
    const int b = blockIdx.x;
    if (!box_has_target(tl, b)) return;  // uniform
    const uint64_t key = tl.key[b];
    const int ix = (int)compact21(key), iy = (int)compact21(key >> 1), iz = (int)compact21(key >> 2);
    const int px = ix >> 1, py = iy >> 1, pz = iz >> 1;
    const int nside = 2 * ring + 1, ncand = nside * nside * nside;
    const int pmax = (1 << (l - 1)) - 1;
    const int wside = 2 * W + 1;
    // 1) parent-level near boxes and their child ranges
    for (int t = threadIdx.x; t < ncand; t += blockDim.x) {
        const int qx = px + t % nside - ring, qy = py + (t / nside) % nside - ring,
                  qz = pz + t / (nside * nside) - ring;
        int first = 0, cnt = 0;
        if (qx >= 0 && qy >= 0 && qz >= 0 && qx <= pmax && qy <= pmax && qz <= pmax) {
            const uint64_t pk = spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2;
            const int pb = find_key(pl.key, pl.nb, pk);
            if (pb >= 0) {
                first = pl.cfirst[pb];
                cnt = (pb + 1 < pl.nb ? pl.cfirst[pb + 1] : tl.nb) - first;
            }
        }
        pfirst[t] = first;
        pcnt[t] = cnt;
    }
    __syncthreads();
    // 2) far children, 8 parents x 8 child slots per wave-64 step
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int k = 0;
        for (int t0 = 0; t0 < ncand; t0 += 8) {
            const int t = t0 + (lane >> 3), cslot = lane & 7;
            bool keep = false;
            int c = 0, off = 0;
            if (t < ncand && cslot < pcnt[t]) {
                c = pfirst[t] + cslot;
                const uint64_t sk = tl.key[c];
                const int dx = ix - (int)compact21(sk), dy = iy - (int)compact21(sk >> 1),
                          dz = iz - (int)compact21(sk >> 2);
                keep = !(abs(dx) <= ring && abs(dy) <= ring && abs(dz) <= ring);
                off = (dx + W) + wside * ((dy + W) + wside * (dz + W));
            }
            const unsigned long long m = __ballot(keep);
            if (keep) {
                const int pos = k + __popcll(m & ((1ULL << lane) - 1ULL));
                src[pos] = c;
                srcoff[pos] = off;
            }
            k += __popcll(m);
        }
        if (lane == 0) {
            nsrc = k;
            atomicAdd(stats, (unsigned long long)k);
        }
    }
    __syncthreads();
    const int ns = nsrc;
    const int me = threadIdx.x;
    const int2 tk0 = me < ntask ? task[me] : make_int2(0, 0);
    const int2 tk1 = me + kFmmThreads < ntask ? task[me + kFmmThreads] : make_int2(0, 0);
    float acc0 = 0.0f, acc1 = 0.0f;
    for (int base = 0; base < ns; base += kM2LBatch) {
        const int nb = min(kM2LBatch, ns - base);
        __syncthreads();
        for (int t = threadIdx.x; t < kM2LBatch * nc; t += blockDim.x) {
            const int s = t / nc, k = t % nc;
            if (s < nb) {
                Ms[s][k] = tl.M[(size_t)src[base + s] * tl.ld + k];
                Ts[s][k] = Ttab[(size_t)srcoff[base + s] * nc + k];
            } else {
                Ms[s][k] = 0.0f;
                Ts[s][k] = 0.0f;
            }
        }
        __syncthreads();
        for (int q = tk0.x; q < tk0.y; q++) {
            const int a = pa[q], g = pg[q];
            float v = 0.0f;
#pragma unroll
            for (int s = 0; s < kM2LBatch; s++) v = fmaf(Ts[s][g], Ms[s][a], v);
            acc0 = fmaf(pw[q], v, acc0);
        }
        for (int q = tk1.x; q < tk1.y; q++) {
            const int a = pa[q], g = pg[q];
            float v = 0.0f;
#pragma unroll
            for (int s = 0; s < kM2LBatch; s++) v = fmaf(Ts[s][g], Ms[s][a], v);
            acc1 = fmaf(pw[q], v, acc1);
        }
    }
    part[me] = acc0;
    part[me + kFmmThreads] = acc1;
    __syncthreads();
    for (int be = threadIdx.x; be < nc; be += blockDim.x) {
        const int2 bt = beta_tasks[be];
        float s = 0.0f;
        for (int t = bt.x; t < bt.y; t++) s += part[t];
        tl.Lc[(size_t)b * tl.ld + be] += s;
    }
}

// ------------------------------------------------- M2L, unrolled per lane ----
//
// With Mhat_a = (-1)^|a| M_a / a!, D_g = g! T_g and Lhat_b = b! L_b the M2L
// contraction is Lhat_b = sum_{|a|+|b|<=p} D_{a+b} Mhat_a: no weights, and
// the (a, b, a+b) index triples are compile-time constants.  A lane owns one
// target box and walks the far-band offsets in the same order as every other
// lane of its wave, so D_{a+b}(offset) is a uniform (scalar) operand and the
// multiply-adds unroll into straight-line code; only Mhat of the source box is
// a per-lane (vector) load.  The outputs are split into passes by the degree of
// b so that the accumulators fit in registers.  The expansion order is chosen
// per Chebyshev shell of the offset (3, 4, 5 box widths).

struct PairIdx {
    short ia, ib, ig;
};

template <int N>
struct PairArr {
    PairIdx v[N > 0 ? N : 1];
};

constexpr int count_pairs(int P, int BLO, int BHI) {
    int k = 0;
    for (int na = 0; na <= P; na++)
        for (int nb = BLO; nb <= BHI; nb++)
            if (na + nb <= P) k += (na + 1) * (na + 2) / 2 * ((nb + 1) * (nb + 2) / 2);
    return k;
}

// Multi-index of each coefficient index (degree-major order), up to degree 12.
struct Tri {
    short x, y, z;
};
struct TriArr {
    Tri v[kMaxNc];
};
constexpr TriArr make_multi() {
    TriArr t{};
    int k = 0;
    for (int n = 0; n <= kMaxP; n++)
        for (int a = n; a >= 0; a--)
            for (int b = n - a; b >= 0; b--) t.v[k++] = {(short)a, (short)b, (short)(n - a - b)};
    return t;
}
constexpr TriArr kMulti = make_multi();

// Term order: blocks of kAlphaBlock source coefficients (held in registers),
// inside a block by the index g = a + b of the translation coefficient D_g,
// so that each D_g (a wave-uniform scalar) is live only across its own terms.
constexpr int kAlphaBlock = 32;
template <int P, int BLO, int BHI>
constexpr PairArr<count_pairs(P, BLO, BHI)> make_pairs() {
    PairArr<count_pairs(P, BLO, BHI)> out{};
    int k = 0;
    const int nca = P >= BLO ? dstart(P - BLO + 1) : 0;  // sources with |a| <= P - BLO
    const int ncg = dstart(P + 1);
    for (int a0 = 0; a0 < nca; a0 += kAlphaBlock) {
        const int a1 = a0 + kAlphaBlock < nca ? a0 + kAlphaBlock : nca;
        for (int ig = 0; ig < ncg; ig++) {
            const Tri g = kMulti.v[ig];
            for (int ia = a0; ia < a1; ia++) {
                const Tri a = kMulti.v[ia];
                const int bx = g.x - a.x, by = g.y - a.y, bz = g.z - a.z;
                if (bx < 0 || by < 0 || bz < 0) continue;
                const int nb = bx + by + bz;
                if (nb < BLO || nb > BHI) continue;
                out.v[k].ia = (short)ia;
                out.v[k].ib = (short)(cidx(bx, by, bz) - dstart(BLO));
                out.v[k].ig = (short)ig;
                k++;
            }
        }
    }
    return out;
}

template <int P, int BLO, int BHI>
struct PairTable {
    static constexpr int n = count_pairs(P, BLO, BHI);
    static constexpr PairArr<n> t = make_pairs<P, BLO, BHI>();
};

constexpr int kTermChunk = 1024;  // terms per fold expression (compiler nesting limit)

template <int P, int BLO, int BHI, int OFF, size_t... I>
__device__ __forceinline__ void m2l_terms(float* acc, const float* __restrict__ Ms,
                                          const float* __restrict__ D, std::index_sequence<I...>) {
    using T = PairTable<P, BLO, BHI>;
    ((acc[T::t.v[OFF + I].ib] =
          fmaf(D[T::t.v[OFF + I].ig], Ms[T::t.v[OFF + I].ia], acc[T::t.v[OFF + I].ib])),
     ...);
}

template <int P, int BLO, int BHI, size_t... C>
__device__ __forceinline__ void m2l_chunks(float* acc, const float* __restrict__ Ms,
                                           const float* __restrict__ D, std::index_sequence<C...>) {
    constexpr int n = PairTable<P, BLO, BHI>::n;
    (m2l_terms<P, BLO, BHI, (int)C * kTermChunk>(
         acc, Ms, D,
         std::make_index_sequence<((int)C + 1) * kTermChunk <= n ? kTermChunk
                                                                 : n - (int)C * kTermChunk>{}),
     ...);
}

template <int P, int BLO, int BHI>
__device__ __forceinline__ void m2l_apply(float* acc, const float* __restrict__ Ms,
                                          const float* __restrict__ D) {
    if constexpr (P >= BLO) {
        constexpr int hi = BHI < P ? BHI : P;
        constexpr int n = PairTable<P, BLO, hi>::n;
        m2l_chunks<P, BLO, hi>(acc, Ms, D,
                               std::make_index_sequence<(n + kTermChunk - 1) / kTermChunk>{});
    }
}

struct M2LLevel {
    int l, nb;
    long gbase;          // index of the level's first box in the all-level box numbering
    const uint64_t* key;
    const float* Mhat;   // [box][nc]
    float* Lc;           // [box][nc]
    const float* D;      // [offset][nc], offsets (dx+5) + 11 (dy+5) + 121 (dz+5)
    const int* grid;     // dense box index grid over the occupied cells
    int lo[3], dim[3];
};

// All M2L levels of one evaluation.  The boxes of levels >= 2 are grouped by
// (level, octant in the parent) with Morton order kept inside a group; a tile
// is up to 64 boxes of one group, so all lanes of a wave share the child
// octant and with it the list of valid far offsets (875 for a near field of
// two cells: the offsets whose source parent is within two cells of the
// target parent).  Work unit = (tile, chunk of that offset list); the lists
// are sorted by Chebyshev shell (3, 4, 5) and cut into chunks of equal cost
// per pass.
constexpr int kM2LMaxLev = 20;
struct M2LPlan {
    int nlev, ntile, nchunk, nc, ld;  // ld: row stride of the expansions
    int nol, s4, s5;                  // offsets per octant list; first entries of shells 4 and 5
    const int4* tiles;                // {first sorted box, count, level index, octant}
    const int* sorted;                // box index within its level, grouped by (level, octant)
    M2LLevel lv[kM2LMaxLev];          // level index li = level - 2
};

// Grouping of the boxes of levels >= 2 by (level, octant): sort keys.
struct LevelOffs {
    int nl;
    int off[kM2LMaxLev + 1];  // level index li covers [off[li], off[li+1])
};

// Boxes without targets get the key kNoGroup and sort after every group.
constexpr int kNoGroup = 255;  // > 8 kM2LMaxLev, within the 8 sorted bits
static_assert(8 * kM2LMaxLev < kNoGroup, "group keys must stay below the sentinel");

__global__ void k_fmm_oct_keys(const uint64_t* __restrict__ bkey, LevelOffs lo, const int* __restrict__ bstart,
                               const int* __restrict__ bend, const int* __restrict__ tpre,
                               unsigned short* __restrict__ okey, int* __restrict__ oval) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= lo.off[lo.nl]) return;
    int li = 0;
    while (li + 1 < lo.nl && lo.off[li + 1] <= i) li++;
    const bool tgt = !tpre || tpre[bend[i]] > tpre[bstart[i]];
    okey[i] = (unsigned short)(tgt ? li * 8 + (int)(bkey[i] & 7) : kNoGroup);
    oval[i] = i - lo.off[li];
}

__global__ void k_fmm_seg_bounds(const unsigned short* __restrict__ sk, int n, int nseg, int* __restrict__ start,
                                 int* __restrict__ end) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = sk[i];
    if (k >= nseg) return;  // boxes without targets
    if (i == 0 || sk[i - 1] != k) start[k] = i;
    if (i == n - 1 || sk[i + 1] != k) end[k] = i + 1;
}

// Tiles of at most 64 boxes per (level, octant) group; entries up to `ub` that
// are not needed get count 0 (their workgroups return at once).
__global__ void k_fmm_tile_table(const int* __restrict__ start, const int* __restrict__ end, int nseg,
                                 int ub, int4* __restrict__ tiles) {
    if (threadIdx.x != 0) return;
    int t = 0;
    for (int sg = 0; sg < nseg; sg++) {
        const int cnt = end[sg] - start[sg];
        for (int k = 0; k < cnt && t < ub; k += 64) tiles[t++] = make_int4(start[sg] + k, min(64, cnt - k), sg >> 3, sg & 7);
    }
    for (; t < ub; t++) tiles[t] = make_int4(0, 0, 0, 0);
}

// The four output-degree ranges computed by separate kernels (registers).
constexpr int kM2LPasses = 4;
constexpr int kPassLo[kM2LPasses] = {0, 5, 7, 9};
constexpr int kPassHi[kM2LPasses] = {4, 6, 8, 10};

constexpr int pass_pairs(int P, int lo, int hi) { return P >= lo ? count_pairs(P, lo, hi < P ? hi : P) : 0; }

template <int P3, int P4, int P5, int BLO, int BHI>
__global__ __launch_bounds__(64, BLO == 0 ? 4 : (BLO == 7 ? 3 : 1)) void k_fmm_m2l_tiled(const M2LPlan pl, const short* __restrict__ olist,
                                                      const int* __restrict__ cbound,
                                                      const float* __restrict__ lscale,
                                                      const float* __restrict__ zeros,
                                                      float* __restrict__ part, size_t part_stride,
                                                      unsigned long long* __restrict__ stats) {
    constexpr int NB = dstart(BHI + 1) - dstart(BLO);
    constexpr bool w3 = pass_pairs(P3, BLO, BHI) > 0, w4 = pass_pairs(P4, BLO, BHI) > 0,
                   w5 = pass_pairs(P5, BLO, BHI) > 0;
    // Blocks are dealt round-robin to the 8 XCDs: renumber so that each XCD
    // runs a contiguous range of tiles (neighbouring tiles share source boxes
    // in that XCD's L2).
    const int nblk = pl.ntile * pl.nchunk;
    const int per = (nblk + 7) >> 3;
    const int w = blockIdx.x;
    const int n = (w & 7) * per + (w >> 3);
    if (n >= nblk) return;
    const int tile = n / pl.nchunk, chunk = n % pl.nchunk;
    const int4 tl = pl.tiles[tile];
    if (tl.y == 0) return;
    const M2LLevel& lv = pl.lv[tl.z];
    const short* ol = olist + tl.w * pl.nol;
    const int nc = pl.nc;
    float acc[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) acc[k] = 0.0f;
    const bool valid = (int)threadIdx.x < tl.y;
    const int b = valid ? pl.sorted[tl.x + threadIdx.x] : 0;
    const uint64_t key = lv.key[b];
    const int ix = (int)compact21(key), iy = (int)compact21(key >> 1), iz = (int)compact21(key >> 2);
    const int cmax = (1 << lv.l) - 1;
    int used[3] = {0, 0, 0};
    const int i1 = cbound[chunk + 1];
    for (int i = cbound[chunk]; i < i1; i++) {
        const int shell = i < pl.s4 ? 3 : (i < pl.s5 ? 4 : 5);  // uniform
        if ((shell == 3 && !w3) || (shell == 4 && !w4) || (shell == 5 && !w5)) continue;
        const int o = ol[i];
        const int dx = o % 11 - 5, dy = (o / 11) % 11 - 5, dz = o / 121 - 5;
        const int sx = ix - dx, sy = iy - dy, sz = iz - dz;
        // the octant's list holds only offsets whose source parent is near the
        // target parent; what remains is the edge of the level
        const bool ok = valid && sx >= 0 && sy >= 0 && sz >= 0 && sx <= cmax && sy <= cmax && sz <= cmax;
        int s = -1;
        if (ok) {
            const int gx = sx - lv.lo[0], gy = sy - lv.lo[1], gz = sz - lv.lo[2];
            if (gx >= 0 && gy >= 0 && gz >= 0 && gx < lv.dim[0] && gy < lv.dim[1] && gz < lv.dim[2])
                s = lv.grid[gx + lv.dim[0] * (gy + lv.dim[1] * gz)];
        }
        used[shell - 3] += s >= 0;
        const float* Ms = s >= 0 ? lv.Mhat + (size_t)s * pl.ld : zeros;
        const float* D = lv.D + (size_t)o * nc;
        if (shell == 3) {
            if constexpr (w3) m2l_apply<P3, BLO, BHI>(acc, Ms, D);
        } else if (shell == 4) {
            if constexpr (w4) m2l_apply<P4, BLO, BHI>(acc, Ms, D);
        } else {
            if constexpr (w5) m2l_apply<P5, BLO, BHI>(acc, Ms, D);
        }
    }
    if (valid) {
        if (pl.nchunk == 1) {
            float* Lb = lv.Lc + (size_t)b * pl.ld + dstart(BLO);
            const float* sc = lscale + dstart(BLO);
#pragma unroll
            for (int k = 0; k < NB; k++) Lb[k] += acc[k] * sc[k];
        } else {
            float* Pb = part + chunk * part_stride + (size_t)(lv.gbase + b) * pl.ld + dstart(BLO);
#pragma unroll
            for (int k = 0; k < NB; k++) Pb[k] = acc[k];
        }
    }
    if (BLO == 0 && stats) {
        // one count per wave and shell: sum of the lanes' source counts
        for (int k = 0; k < 3; k++) {
            int u = used[k];
            for (int off = 32; off > 0; off >>= 1) u += __shfl_xor(u, off);
            if (threadIdx.x == 0 && u) {
                atomicAdd(stats, (unsigned long long)u);
                atomicAdd(stats + 2 + k, (unsigned long long)u);
            }
        }
    }
}

// L[box][k] += lscale[k] * sum over chunks of the partial sums (fixed order);
// rows of stride ld (a multiple of 4), the padding past nc left alone.  Each
// lane takes four consecutive coefficients of one row.
// Rows of boxes without targets were not written by the M2L and are skipped.
__global__ __launch_bounds__(256) void k_fmm_m2l_reduce(const float* __restrict__ part, size_t stride, int nchunk,
                                                        int nc, int ld, const float* __restrict__ lscale,
                                                        float* __restrict__ Lc, size_t n4,
                                                        const int* __restrict__ bstart, const int* __restrict__ bend,
                                                        const int* __restrict__ tpre) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n4) return;
    const size_t e = 4 * q;
    const size_t row = e / (size_t)ld;
    const int k = (int)(e - row * (size_t)ld);
    if (k >= nc) return;
    if (tpre && tpre[bend[row]] <= tpre[bstart[row]]) return;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (int c = 0; c < nchunk; c++) {
        const float4 v = *reinterpret_cast<const float4*>(part + c * stride + e);
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.w += v.w;
    }
    float4 l = *reinterpret_cast<float4*>(Lc + e);
    l.x += s.x * lscale[k];
    if (k + 1 < nc) l.y += s.y * lscale[k + 1];
    if (k + 2 < nc) l.z += s.z * lscale[k + 2];
    if (k + 3 < nc) l.w += s.w * lscale[k + 3];
    *reinterpret_cast<float4*>(Lc + e) = l;
}

// Offsets of shells 3..5 (ring 2) sorted by shell; chunk bounds per pass such
// that each chunk carries about the same number of multiply-adds.
static void m2l_offsets(std::vector<short>& ol, int& nol, int& s4, int& s5) {
    // per child octant c: offsets o (source = target - o) of shells 3..5 whose
    // source parent, floor((c - o) / 2) per axis, is within two cells
    ol.clear();
    for (int c = 0; c < 8; c++) {
        const int cc[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
        const size_t base = ol.size();
        for (int shell = 3; shell <= 5; shell++) {
            if (shell == 4) s4 = (int)(ol.size() - base);
            if (shell == 5) s5 = (int)(ol.size() - base);
            for (int o = 0; o < 1331; o++) {
                const int d[3] = {o % 11 - 5, (o / 11) % 11 - 5, o / 121 - 5};
                if (std::max(std::abs(d[0]), std::max(std::abs(d[1]), std::abs(d[2]))) != shell) continue;
                bool ok = true;
                for (int a = 0; a < 3; a++) {
                    const int q = cc[a] - d[a];
                    const int pq = q >= 0 ? q / 2 : -((-q + 1) / 2);  // floor(q / 2)
                    ok = ok && pq >= -2 && pq <= 2;
                }
                if (ok) ol.push_back((short)o);
            }
        }
        nol = (int)(ol.size() - base);  // the same for every octant (875)
    }
}

static std::vector<int> m2l_chunk_bounds(int p3, int p4, int p5, int S, int s4, int s5, int nol) {
    std::vector<int> cb;
    for (int ps = 0; ps < kM2LPasses; ps++) {
        const int c3 = pass_pairs(p3, kPassLo[ps], kPassHi[ps]);
        const int c4 = pass_pairs(p4, kPassLo[ps], kPassHi[ps]);
        const int c5 = pass_pairs(p5, kPassLo[ps], kPassHi[ps]);
        auto cost = [&](int i) {
            const int c = i < s4 ? c3 : (i < s5 ? c4 : c5);
            return c ? (double)c + 64.0 : 0.0;  // + box lookup and loads
        };
        double tot = 0;
        for (int i = 0; i < nol; i++) tot += cost(i);
        cb.push_back(0);
        double run = 0;
        int c = 1;
        for (int i = 0; i < nol && c < S; i++) {
            run += cost(i);
            while (c < S && run >= tot * c / S) {
                cb.push_back(i + 1);
                c++;
            }
        }
        while ((int)cb.size() < (ps + 1) * (S + 1) - 1) cb.push_back(nol);
        cb.push_back(nol);
    }
    return cb;
}

// Estimated ms of the level-dependent passes of a leaf level on MI355X,
// fitted to measured pass times of the unrolled M2L (p = 10, ring 2; 16 plume
// states and levels, within 8%) and of the leaf kernel: per box of a level, a
// far-band offset costs 0.022 ns when its source cell is empty and 0.27 ns
// more when it holds a box (the pass is bound by the source expansion loads,
// so the cost scales with the coefficient count), and the upward and downward
// passes 7 ns; a near-field particle pair costs 0.34 ps and a leaf box 11 ns
// of leaf-kernel set-up (near-field run list, L2P; within 11%).
static double fmm_m2l_cost_ms(int nc, int nbox, double far_pairs) {
    const double s = nc / 286.0;
    return s * (2.16e-8 * 875.0 * nbox + 2.7e-7 * far_pairs) + 7.0e-6 * nbox;
}
static double fmm_leaf_cost_ms(int nbox, double p2p_pairs) { return 3.4e-10 * p2p_pairs + 1.13e-5 * nbox; }

// M2L orders on the Chebyshev shells 3, 4, 5 of the far band (near field of
// two cells), fitted to the FP64 restatement on cones: p = 10 keeps its error
// with (10, 8, 7); lower orders lose more per dropped degree, so p = 8 uses
// (8, 7, 6), p = 6 uses (6, 5, 4) and p = 4 keeps 4 throughout.  `full`
// (rotational images, whose fields cancel near the axis) keeps p on every
// shell.  Zeros: no compiled term list (the table-driven M2L is used).
static void m2l_orders(int p, int ring, bool full, int& p3, int& p4, int& p5) {
    p3 = p4 = p5 = 0;
    if (ring != 2) return;
    if (p == 10) {
        p3 = 10;
        p4 = full ? 10 : 8;
        p5 = full ? 10 : 7;
    } else if (p == 8) {
        p3 = 8;
        p4 = 7;
        p5 = 6;
    } else if (p == 6) {
        p3 = 6;
        p4 = 5;
        p5 = 4;
    } else if (p == 4) {
        p3 = p4 = p5 = 4;
    }
}

template <int P3, int P4, int P5>
static void launch_m2l_passes(const M2LPlan& pl, const short* olist, const int* cbound,
                              const float* lscale, const float* zeros, float* part, size_t stride,
                              unsigned long long* stats, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const int nblk = pl.ntile * pl.nchunk;
    const dim3 g(8 * ((nblk + 7) / 8)), b(64);
    const int S1 = pl.nchunk + 1;
    (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL((k_fmm_m2l_tiled<P3, P4, P5, 0, 4>), g, b, 0, st, pl, olist, cbound, lscale,
                       zeros, part, stride, stats);
    (void)hipEventRecord(e1, st);
    if constexpr (P3 >= 5)
        hipLaunchKernelGGL((k_fmm_m2l_tiled<P3, P4, P5, 5, 6>), g, b, 0, st, pl, olist, cbound + S1,
                           lscale, zeros, part, stride, stats);
    if constexpr (P3 >= 7)
        hipLaunchKernelGGL((k_fmm_m2l_tiled<P3, P4, P5, 7, 8>), g, b, 0, st, pl, olist, cbound + 2 * S1,
                           lscale, zeros, part, stride, stats);
    if constexpr (P3 >= 9)
        hipLaunchKernelGGL((k_fmm_m2l_tiled<P3, P4, P5, 9, 10>), g, b, 0, st, pl, olist, cbound + 3 * S1,
                           lscale, zeros, part, stride, stats);
}

static void launch_m2l_tiled(const M2LPlan& pl, int p3, int p4, const FmmDevTables* tb, const int* cbound,
                             float* part, size_t stride, unsigned long long* stats, hipStream_t st,
                             hipEvent_t e0, hipEvent_t e1) {
    if (p3 == 10 && p4 == 10)
        launch_m2l_passes<10, 10, 10>(pl, tb->olist, cbound, tb->lscale, tb->zeros, part, stride, stats, st, e0, e1);
    else if (p3 == 10)
        launch_m2l_passes<10, 8, 7>(pl, tb->olist, cbound, tb->lscale, tb->zeros, part, stride, stats, st, e0, e1);
    else if (p3 == 8)
        launch_m2l_passes<8, 7, 6>(pl, tb->olist, cbound, tb->lscale, tb->zeros, part, stride, stats, st, e0, e1);
    else if (p3 == 6)
        launch_m2l_passes<6, 5, 4>(pl, tb->olist, cbound, tb->lscale, tb->zeros, part, stride, stats, st, e0, e1);
    else if (p3 == 4)
        launch_m2l_passes<4, 4, 4>(pl, tb->olist, cbound, tb->lscale, tb->zeros, part, stride, stats, st, e0, e1);
}

// Offset list and chunk bounds of the tiled M2L, made once per (order, chunks).
static int m2l_bounds(FmmDevTables* tb, int p3, int p4, int p5, int S, const int** out) {
    if (!tb->olist) {
        std::vector<short> ol;
        int nol = 0, s4 = 0, s5 = 0;
        m2l_offsets(ol, nol, s4, s5);
        if (upload(&tb->olist, ol) != SHP_OK) return SHP_ERR_HIP;
        tb->olist_n = nol;
        tb->olist_s4 = s4;
        tb->olist_s5 = s5;
    }
    const int key = ((p3 * 16 + p4) * 16 + p5) * 128 + S;
    if (tb->cbound_key != key) {
        const std::vector<int> cb = m2l_chunk_bounds(p3, p4, p5, S, tb->olist_s4, tb->olist_s5, tb->olist_n);
        if (upload(&tb->cbound, cb) != SHP_OK) return SHP_ERR_HIP;
        tb->cbound_key = key;
    }
    *out = tb->cbound;
    return SHP_OK;
}

// v[row][k] *= sc[k] over rows of stride ld (padding past nc left alone):
// one workgroup per row, one lane per coefficient.
__global__ void k_fmm_scale(float* __restrict__ v, const float* __restrict__ sc, int nc, int ld) {
    const int k = threadIdx.x;
    if (k < nc) v[(size_t)blockIdx.x * ld + k] *= sc[k];
}

// Dense cell grids of the levels 0 .. nl-1 in one launch: box b of the
// concatenated level key arrays (level l holds boxes [boff[l], boff[l+1]))
// writes its index within its level into that level's grid.
constexpr int kGridLevels = 24;
struct GridScatter {
    const uint64_t* key;  // all levels' keys, level after level
    int nl;
    int boff[kGridLevels + 1];
    int lo[kGridLevels][3], dim[kGridLevels][3];
    long long goff[kGridLevels];
};
__global__ void k_fmm_grid_scatter(const GridScatter S, int* __restrict__ grid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= S.boff[S.nl]) return;
    int l = 0;
    while (i >= S.boff[l + 1]) l++;
    const uint64_t k = S.key[i];
    const int gx = (int)compact21(k) - S.lo[l][0], gy = (int)compact21(k >> 1) - S.lo[l][1],
              gz = (int)compact21(k >> 2) - S.lo[l][2];
    if (gx >= 0 && gy >= 0 && gz >= 0 && gx < S.dim[l][0] && gy < S.dim[l][1] && gz < S.dim[l][2])
        grid[S.goff[l] + gx + (long long)S.dim[l][0] * (gy + (long long)S.dim[l][1] * gz)] = i - S.boff[l];
}

// ------------------------------------------------------------------- L2L ----

// L2L with the same octant factorisation: split the parent's sum by the parity
// class of b,
//   Q_r[k] = 1/2 sum_{b>=k, class(b)=r} C(b,k) 2^-|b| L_b,
// then each child takes L'_k += s(o, k) sum_r s(o, r) Q_r[k] (an inverse
// transform over r per k).  One wave per parent; flat per-lane term lists.
__global__ __launch_bounds__(kShiftLanes) void k_fmm_l2l(Level par, Level ch, int nc,
                                                         const int4* __restrict__ mi,
                                                         const int2* __restrict__ term, int nt) {
    extern __shared__ float sm[];
    float* Lp = sm;      // [nc]
    float* Q = sm + nc;  // [8][nc]: by parity class, then by child octant
    const int b = blockIdx.x, t = threadIdx.x;
    if (!box_has_target(par, b)) return;  // uniform: no child holds a target
    const int c0 = par.cfirst[b];
    const int c1 = b + 1 < par.nb ? par.cfirst[b + 1] : ch.nb;
    const float* Lsrc = par.Lc + (size_t)b * par.ld;
    for (int k = t; k < nc; k += kShiftLanes) Lp[k] = Lsrc[k];
    for (int i = t; i < 8 * nc; i += kShiftLanes) Q[i] = 0.0f;
    __syncthreads();
    float acc = 0.0f;
    for (int j0 = 0; j0 < nt; j0 += kShiftUnroll) {
        int2 e[kShiftUnroll];
#pragma unroll
        for (int u = 0; u < kShiftUnroll; u++) e[u] = term[(j0 + u) * kShiftLanes + t];
#pragma unroll
        for (int u = 0; u < kShiftUnroll; u++) {
            acc = fmaf(__int_as_float(e[u].y), Lp[e[u].x & 0xfff], acc);
            if (e[u].x >> 24) {
                Q[(e[u].x >> 12) & 0xfff] = acc;
                acc = 0.0f;
            }
        }
    }
    __syncthreads();
    for (int k = t; k < nc; k += kShiftLanes) {
        float u[8];
#pragma unroll
        for (int o = 0; o < 8; o++) u[o] = Q[o * nc + k];
#pragma unroll
        for (int d = 0; d < 3; d++)
#pragma unroll
            for (int o = 0; o < 8; o++)
                if (!(o & (1 << d))) {
                    const float x = u[o], y = u[o | (1 << d)];
                    u[o] = x - y;             // child bit 0: s = (-1)^class bit
                    u[o | (1 << d)] = x + y;  // child bit 1: s = +1
                }
        const int4 K = mi[k];
#pragma unroll
        for (int o = 0; o < 8; o++) {
            const int par_k = ((o & 1) ? 0 : K.x) + ((o & 2) ? 0 : K.y) + ((o & 4) ? 0 : K.z);
            Q[o * nc + k] = (par_k & 1) ? -u[o] : u[o];
        }
    }
    __syncthreads();
    for (int c = c0; c < c1; c++) {
        const int o = (int)(ch.key[c] & 7);
        float* Lc = ch.Lc + (size_t)c * ch.ld;
        for (int k = t; k < nc; k += kShiftLanes) Lc[k] += Q[o * nc + k];
    }
}

__global__ void k_fmm_zero(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

// ------------------------------------------------------------ L2P + P2P ----

__device__ __forceinline__ void fmm_two_sum(float& hi, float& lo, float p) {
    const float s = hi + p;
    const float bp = s - hi;
    lo += (hi - (s - bp)) + (p - bp);
    hi = s;
}

// Dense cell -> box index grid of one level (nullptr: search the sorted keys).
struct GridView {
    const int* g;
    int lo[3], dim[3];
};

__device__ __forceinline__ int find_cell(const Level& lv, const GridView& gv, int qx, int qy, int qz) {
    if (gv.g) {
        const int gx = qx - gv.lo[0], gy = qy - gv.lo[1], gz = qz - gv.lo[2];
        if (gx < 0 || gy < 0 || gz < 0 || gx >= gv.dim[0] || gy >= gv.dim[1] || gz >= gv.dim[2]) return -1;
        return gv.g[gx + gv.dim[0] * (gy + gv.dim[1] * gz)];
    }
    return find_key(lv.key, lv.nb, spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2);
}

// Work of a uniform leaf level l, counted over every stride-th box b of the
// level (one wave per box at a time, its lanes splitting the cell lookups;
// the waves of the grid stride over the sampled boxes): out[0] += n_b x
// (particles in the near cells of b) = P2P pairs; out[1] += far-band boxes of
// b = occupied children of the parent's near boxes minus the occupied near
// boxes of b = M2L box pairs.  One pair of atomics per workgroup.
constexpr int kLevelCostBlocks = 512;
constexpr int kLevelCostMax = 4;  // candidate levels counted by one launch (blockIdx.y)
struct LevelCostArgs {
    Level lv[kLevelCostMax], par[kLevelCostMax];
    GridView g[kLevelCostMax], gp[kLevelCostMax];
    int stride[kLevelCostMax];
    unsigned long long* out[kLevelCostMax];
};
__global__ __launch_bounds__(256) void k_fmm_level_cost(const LevelCostArgs A, int ring) {
    const Level& lv = A.lv[blockIdx.y];
    const Level& par = A.par[blockIdx.y];
    const GridView& g = A.g[blockIdx.y];
    const GridView& gp = A.gp[blockIdx.y];
    const int stride = A.stride[blockIdx.y];
    unsigned long long* out = A.out[blockIdx.y];
    __shared__ long red[2][4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int side = 2 * ring + 1, ncell = side * side * side;
    long nsum_all = 0, far = 0;
    for (long s = (long)blockIdx.x * 4 + wv; s * stride < lv.nb; s += (long)gridDim.x * 4) {
        const long b = s * stride;
        if (!box_has_target(lv, (int)b)) continue;  // wave-uniform
        const uint64_t key = lv.key[b];
        const int ix = (int)compact21(key), iy = (int)compact21(key >> 1), iz = (int)compact21(key >> 2);
        long nsum = 0;
        for (int t = lane; t < ncell; t += 64) {
            const int dx = t % side - ring, dy = (t / side) % side - ring, dz = t / (side * side) - ring;
            const int c = find_cell(lv, g, ix + dx, iy + dy, iz + dz);
            if (c >= 0) {
                nsum += lv.end[c] - lv.start[c];
                far -= 1;
            }
            const int q = find_cell(par, gp, (ix >> 1) + dx, (iy >> 1) + dy, (iz >> 1) + dz);
            if (q >= 0) far += (q + 1 < par.nb ? par.cfirst[q + 1] : lv.nb) - par.cfirst[q];
        }
        nsum_all += nsum * (lv.end[b] - lv.start[b]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        nsum_all += __shfl_xor(nsum_all, off);
        far += __shfl_xor(far, off);
    }
    if (lane == 0) {
        red[0][wv] = nsum_all;
        red[1][wv] = far;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const long a = red[0][0] + red[0][1] + red[0][2] + red[0][3];
        const long f = red[1][0] + red[1][1] + red[1][2] + red[1][3];
        if (a || f) {
            atomicAdd(out, (unsigned long long)a);
            atomicAdd(out + 1, (unsigned long long)f);
        }
    }
}

// Leaf work units: a leaf's targets in runs of kLeafUnit, one wave each, so
// that the dense leaves of a plume core spread over many waves.
constexpr int kLeafUnit = 64;
constexpr int kLeafTile = 256;  // sources staged in LDS per barrier pair
constexpr int kLeafXcdRun = 16;  // consecutive leaf units dealt to one XCD

// Sorted ion i relative to the centre of its own leaf in units of the leaf
// half-width h (FP32), with its charge.  A source j of leaf S seen from leaf T
// is then spf[j] + 2 (cell(S) - cell(T)): the shift is a small integer, exact
// in FP32, so staging needs no FP64 arithmetic.
__global__ void k_fmm_leaf_rel(int na, int L, const uint64_t* __restrict__ keys, const double* __restrict__ root,
                               const double4* __restrict__ sp, float4* __restrict__ spf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= na) return;
    double c[3], h;
    box_center(root, L, keys[i] >> (3 * (kMaxLevel - L)), c, h);
    const double ih = 1.0 / h;
    const double4 r = sp[i];
    spf[i] = make_float4((float)((r.x - c[0]) * ih), (float)((r.y - c[1]) * ih), (float)((r.z - c[2]) * ih),
                         (float)r.w);
}

// Per leaf: the number of work units and the number of near-field sources.
__global__ __launch_bounds__(256) void k_fmm_unit_count(Level lv, GridView gv, int L, int ring,
                                                        int* __restrict__ cnt, int* __restrict__ nsrc) {
    // one wave per leaf box, its lanes splitting the near-field cell lookups
    const int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= lv.nb) return;  // wave-uniform
    if (!box_has_target(lv, b)) {  // no unit
        if (lane == 0) {
            cnt[b] = 0;
            nsrc[b] = 0;
        }
        return;
    }
    const uint64_t key = lv.key[b];
    const int ix = (int)compact21(key), iy = (int)compact21(key >> 1), iz = (int)compact21(key >> 2);
    const int cmax = (1 << L) - 1;
    const int nside = 2 * ring + 1, ncand = nside * nside * nside;
    int ns = 0;
    for (int t = lane; t < ncand; t += 64) {
        const int qx = ix + t % nside - ring, qy = iy + (t / nside) % nside - ring, qz = iz + t / (nside * nside) - ring;
        if (qx < 0 || qy < 0 || qz < 0 || qx > cmax || qy > cmax || qz > cmax) continue;
        const int c = find_cell(lv, gv, qx, qy, qz);
        if (c >= 0) ns += lv.end[c] - lv.start[c];
    }
    for (int off = 32; off > 0; off >>= 1) ns += __shfl_xor(ns, off);
    if (lane == 0) {
        cnt[b] = max(1, (lv.end[b] - lv.start[b] + kLeafUnit - 1) / kLeafUnit);
        nsrc[b] = ns;
    }
}

// Sort keys of the units: work = targets x near-field sources (at least 1);
// slots past the unit count get 0 and sort last.
__global__ void k_fmm_unit_keys(const int2* __restrict__ units, const int* __restrict__ nunits, int umax,
                                Level lv, const int* __restrict__ nsrc, unsigned* __restrict__ key,
                                int* __restrict__ val) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= umax) return;
    unsigned w = 0;
    if (u < *nunits) {
        const int2 un = units[u];
        const int nt = min(kLeafUnit, lv.end[un.x] - lv.start[un.x] - kLeafUnit * un.y);
        w = (unsigned)nt * (unsigned)min(max(1, nsrc[un.x]), 1 << 25);  // < 2^31
    }
    key[u] = w;
    val[u] = u;
}

__global__ void k_fmm_unit_fill(const int* __restrict__ cnt, const int* __restrict__ offs, int nb,
                                int2* __restrict__ units, int* __restrict__ nunits) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    for (int k = 0; k < cnt[b]; k++) units[offs[b] + k] = make_int2(b, k);
    if (b == nb - 1) *nunits = offs[b] + cnt[b];
}

// L2P share of lane group g of G: the gradient of sum_b L_b u^b over the
// multi-indices with b_x = g mod G (u in units of the leaf half-width).
// Lsh holds the local coefficients in the order of the loops below: by a,
// then b, then c (coefficient (a, b, c) at l2p_pos(p, a, b, c)).
__device__ __forceinline__ int l2p_pos(int p, int a, int b, int c) {
    const int q = p - a;
    return dstart(p + 1) - dstart(q + 1) + b * (q + 1) - b * (b - 1) / 2 + c;
}

__device__ __forceinline__ void leaf_l2p(const float* __restrict__ Lsh, int p, int g, int G, float ux,
                                         float uy, float uz, float& gx, float& gy, float& gz) {
    gx = gy = gz = 0.0f;
    for (int a = g; a <= p; a += G) {
        float xa = 1.0f, xam1 = 0.0f;
        for (int k = 0; k < a; k++) {
            xam1 = xa;
            xa *= ux;
        }
        const float fa = (float)a;
        int k = dstart(p + 1) - dstart(p - a + 1);  // first coefficient of this a
        float yb = 1.0f, ybm1 = 0.0f, fb = 0.0f;
        for (int bb = 0; a + bb <= p; bb++) {
            float zc = 1.0f, zcm1 = 0.0f, fc = 0.0f;
            for (int cc = 0; a + bb + cc <= p; cc++) {
                const float Lk = Lsh[k++];
                gx = fmaf(Lk * fa, xam1 * yb * zc, gx);
                gy = fmaf(Lk * fb, xa * ybm1 * zc, gy);
                gz = fmaf(Lk * fc, xa * yb * zcm1, gz);
                zcm1 = zc;
                zc *= uz;
                fc += 1.0f;
            }
            ybm1 = yb;
            yb *= uy;
            fb += 1.0f;
        }
    }
}

// One softened pair in leaf units: s += q.w (u - q.xyz) / (|u - q.xyz|^2 + eps2)^(3/2).
__device__ __forceinline__ void leaf_pair(float ux, float uy, float uz, const float4 q, float eps2f,
                                          float& sx, float& sy, float& sz) {
    const float dx = ux - q.x, dy = uy - q.y, dz = uz - q.z;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2f)));
    const float inv = __builtin_amdgcn_rsqf(r2);
    const float w = (q.w * inv) * (inv * inv);
    sx = fmaf(dx, w, sx);
    sy = fmaf(dy, w, sy);
    sz = fmaf(dz, w, sz);
}

// Stages the near-field sources [jb, jb + kLeafTile) relative to the leaf
// centre in units of the half-width (zero charge past nsrc): each source's
// offset from its own leaf centre plus twice the cell offset of its leaf
// (packed per run in rl; rf holds each run's start minus its prefix sum).
// Each lane takes
// kLeafTile / 64 consecutive sources: a search, from its run of the previous
// tile, for the run holding the first (the last r with rpre[r] <= j), then a
// walk forward over the runs.
__device__ __forceinline__ void leaf_stage(float4* tile, int jb, int lane, int nsrc, int nrun,
                                           const int* rpre, const int* rf, const int* rl,
                                           const float4* __restrict__ spf, int& r) {
    __syncthreads();
    const int j0 = jb + (kLeafTile / 64) * lane;
    if (j0 < nsrc) {
        // r (the lane's run from the previous tile, rpre[r] <= j0) moves to the
        // last run with rpre[r] <= j0: doubling steps, then halving ones
        int step = 1;
        while (r + step < nrun && rpre[r + step] <= j0) {
            r += step;
            step <<= 1;
        }
        while (step > 1) {
            step >>= 1;
            if (r + step < nrun && rpre[r + step] <= j0) r += step;
        }
    }
#pragma unroll
    for (int k = 0; k < kLeafTile / 64; k++) {
        const int j = j0 + k;
        float4 sv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (j < nsrc) {
            while (rpre[r + 1] <= j) r++;
            const float4 rs = spf[rf[r] + j];
            const int o = rl[r];  // doubled cell offset of the run, three signed bytes
            sv = make_float4(rs.x + (float)(signed char)(o & 0xff), rs.y + (float)(signed char)((o >> 8) & 0xff),
                             rs.z + (float)(signed char)((o >> 16) & 0xff), rs.w);
        }
        tile[(kLeafTile / 64) * lane + k] = sv;
    }
    __syncthreads();
}

// Acceleration of sorted target i from its sum in leaf units (status 1 only).
__device__ __forceinline__ void leaf_write(int i, double qi, double sx, double sy, double sz,
                                           const int* __restrict__ perm, const int* __restrict__ status,
                                           const int* __restrict__ species, double K, double h,
                                           double* __restrict__ ax, double* __restrict__ ay,
                                           double* __restrict__ az) {
    const int oi = perm[i];
    if (status[oi] == 1) {
        const double kc = K * qi / species_mass(species[oi]) / (h * h);
        ax[oi] = sx * kc;
        ay[oi] = sy * kc;
        az[oi] = sz * kc;
    }
}

// One wave per leaf work unit.  The near-field sources (particles of the
// leaves within `ring` cells) are listed as runs of the sorted particles; the
// unit's targets are then processed in blocks of T lanes, with G = 64 / T lane
// groups that split the sources and the L2P terms, so a leaf of ~20 targets
// keeps ~60 of the 64 lanes busy.  Pairs are FP32 relative to
// the leaf centre (units of h) with compensated per-target sums; the group
// partials are added in FP64, scaled by K q/m and scattered to the caller's
// order.
__global__ __launch_bounds__(64) void k_fmm_leaf(
    Level lv, int L, int ring, GridView gv, const double* __restrict__ root,
    const double4* __restrict__ sp, const int* __restrict__ perm, const int* __restrict__ status,
    const int* __restrict__ species, const float4* __restrict__ spf, const int4* __restrict__ mi, int nc, int p,
    double K,
    const int2* __restrict__ units, const int* __restrict__ uorder, const int* __restrict__ nunits, unsigned long long* __restrict__ stats, double* __restrict__ ax,
    double* __restrict__ ay, double* __restrict__ az) {
    __shared__ float4 tile[kLeafTile + 64];  // staged sources, then 64 zero-charge entries
    __shared__ float Lsh[kMaxNc];
    __shared__ int rf[128], rl[128], rpre[129];
    __shared__ unsigned char rt[128];
    double(*red)[64] = reinterpret_cast<double(*)[64]>(tile);  // group partials, after a block's tiles
    // Units run longest first (uorder sorts them by targets x near-field
    // sources, descending), so the dense units of a plume core do not finish
    // last.
    // Blocks are dealt to the 8 XCDs in turn.  Each run of kLeafXcdRun
    // consecutive units in that order (mostly units of one dense leaf, which
    // share their near field) goes to one XCD, so its L2 fetches the near field
    // once; the longest-first order holds at the granularity of 8 runs.
    const int w = blockIdx.x, grp = w / (8 * kLeafXcdRun), rr = w % (8 * kLeafXcdRun);
    const int un = grp * 8 * kLeafXcdRun + (rr & 7) * kLeafXcdRun + (rr >> 3);
    if (un >= *nunits) return;
    const int2 unit = units[uorder[un]];
    const int b = unit.x, lane = threadIdx.x;
    tile[kLeafTile + lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const uint64_t key = lv.key[b];
    double c[3], h;
    box_center(root, L, key, c, h);
    const double ih = 1.0 / h;
    const float eps2f = (float)(kSoftening / (h * h));
    for (int k = lane; k < nc; k += 64) {
        const int4 A = mi[k];
        Lsh[l2p_pos(p, A.x, A.y, A.z)] = lv.Lc[(size_t)b * lv.ld + k];
    }
    const int ix = (int)compact21(key), iy = (int)compact21(key >> 1), iz = (int)compact21(key >> 2);
    const int nside = 2 * ring + 1, ncand = nside * nside * nside;
    const int cmax = (1 << L) - 1;
    // near-field runs, compacted in candidate order
    int nrun = 0;
    for (int t0 = 0; t0 < ncand; t0 += 64) {
        const int t = t0 + lane;
        int f = 0, e = 0;
        if (t < ncand) {
            const int qx = ix + t % nside - ring, qy = iy + (t / nside) % nside - ring,
                      qz = iz + t / (nside * nside) - ring;
            if (qx >= 0 && qy >= 0 && qz >= 0 && qx <= cmax && qy <= cmax && qz <= cmax) {
                const int nb = find_cell(lv, gv, qx, qy, qz);
                if (nb >= 0) {
                    f = lv.start[nb];
                    e = lv.end[nb];
                }
            }
        }
        const bool has = e > f;
        const unsigned long long m = __ballot(has);
        if (has) {
            const int pos = nrun + __popcll(m & ((1ull << lane) - 1ull));
            rf[pos] = f;
            rl[pos] = e - f;
            rt[pos] = (unsigned char)t;
        }
        nrun += __popcll(m);
    }
    __syncthreads();
    // prefix sums of the run lengths (wave scan)
    int carry = 0;
    for (int r0 = 0; r0 < nrun; r0 += 64) {
        int v = r0 + lane < nrun ? rl[r0 + lane] : 0;
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(v, d);
            if (lane >= d) v += y;
        }
        if (r0 + lane < nrun) rpre[r0 + lane + 1] = carry + v;
        carry += __shfl(v, 63);
    }
    if (lane == 0) rpre[0] = 0;
    __syncthreads();
    // per run, for the staging: rf[r] - rpre[r] (so source j of run r is
    // spf[rf[r] + j]) and, in place of its length, the cell offset doubled as
    // three signed bytes
    for (int r = lane; r < nrun; r += 64) {
        const int t = rt[r];
        const int ox = 2 * (t % nside - ring), oy = 2 * ((t / nside) % nside - ring), oz = 2 * (t / (nside * nside) - ring);
        rl[r] = (ox & 0xff) | ((oy & 0xff) << 8) | ((oz & 0xff) << 16);
        rf[r] -= rpre[r];
    }
    __syncthreads();
    const int nsrc = carry;
    const int s0 = lv.start[b] + kLeafUnit * unit.y, s1 = min(s0 + kLeafUnit, lv.end[b]);
    const int nt = s1 - s0;
    if (lane == 0) atomicAdd(stats + 1, (unsigned long long)nsrc * (unsigned long long)nt);
    for (int tb = s0; tb < s1;) {
        const int rem = s1 - tb;
        // T targets x G source groups per block: 64 targets while more than 53
        // remain, then 32 while more than 32 remain (two source groups), then
        // all that remain (64 / T groups), so every block keeps most lanes busy
        const int T = rem >= 54 ? 64 : (rem > 32 ? 32 : rem), G = 64 / T;
        const int g = lane / T, tl = lane - g * T;
        const bool active = g < G;
        const int i = tb + tl;
        const int tend = min(tb + T, s1);
        tb += T;
        const bool valid = active && i < tend;
        const float4 rv = spf[valid ? i : s0];
        const float ux = rv.x, uy = rv.y, uz = rv.z;
        // L2P share of this group: gradient of sum L_b u^b over b_x = g mod G
        float gx = 0.0f, gy = 0.0f, gz = 0.0f;
        if (active) leaf_l2p(Lsh, p, g, G, ux, uy, uz, gx, gy, gz);
        // far part in units of 1/h^2 relative to the P2P sum expressed in u units:
        // P2P below runs in u = (x - c)/h, so sum q u/|u|^3 = h^2 sum q d/|d|^3.
        float hx = -gx, hy = -gy, hz = -gz, lx = 0.0f, ly = 0.0f, lz = 0.0f;
        int run = 0;  // this lane's run in the staging, carried from tile to tile
        const int nit_full = (kLeafTile + G - 1) / G;  // trip count of a full tile
        for (int jb = 0; jb < nsrc; jb += kLeafTile) {
            leaf_stage(tile, jb, lane, nsrc, nrun, rpre, rf, rl, spf, run);
            // group g takes the contiguous chunk [g nit, (g + 1) nit) of the
            // tile; the trip count is wave-uniform and entries past the staged
            // sources have zero charge
            const int cnt = min(kLeafTile, nsrc - jb);
            const int nit = cnt == kLeafTile ? nit_full : (cnt + G - 1) / G;
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            if (active) {
                const float4* tg = tile + g * nit;
#pragma unroll 8
                for (int it = 0; it < nit; it++) leaf_pair(ux, uy, uz, tg[it], eps2f, sx, sy, sz);
            }
            fmm_two_sum(hx, lx, sx);
            fmm_two_sum(hy, ly, sy);
            fmm_two_sum(hz, lz, sz);
        }
        __syncthreads();  // the last tile is read: its space holds the group partials
        red[0][lane] = (double)hx + (double)lx;
        red[1][lane] = (double)hy + (double)ly;
        red[2][lane] = (double)hz + (double)lz;
        __syncthreads();
        if (g == 0 && valid) {
            double sxd = 0.0, syd = 0.0, szd = 0.0;
            for (int gg = 0; gg < G; gg++) {
                sxd += red[0][gg * T + tl];
                syd += red[1][gg * T + tl];
                szd += red[2][gg * T + tl];
            }
            leaf_write(i, sp[i].w, sxd, syd, szd, perm, status, species, K, h, ax, ay, az);
        }
    }
}

// ------------------------------------------------------------------- M2P ----
// Treecode and outlier evaluation (FmmType treecode / fmmOutliers,
// structures.h:283-288; m2p and m2outlier_p, fmm_cpukernel.cpp:1319-1533,
// 2429-2776): the multipole expansions of the far boxes evaluated at a point.
// For a point X the far boxes of level l are those whose parent is near X's
// cell one level up (at level 0: the root) and that are not near X's own
// cell floor((X - x0) / w_l), not clamped, so X may lie outside the root cube.
// With u = (X - c)/h for box centre c and half size h:
//   E_k = -(1/h^2) sum_{|a|<=p} (-1)^|a| M_a (a_k + 1) T_{a+e_k}(u)
// (M_a raw) = -(1/h^2) sum Mhat_a (a+e_k)! T_{a+e_k}(u) (Mhat_a = (-1)^|a|
// M_a / a!, as the M2L stores them), T from the Taylor recurrence up to
// degree p + 1.
constexpr int kMaxPointLevels = 17;  // levels 0..16
constexpr int kM2PMaxP = 10;         // three degree slices per lane fit 64 KB of LDS
constexpr int kM2PList = 1024;       // candidate boxes of one level: (2 ring + 1)^3 parents x 8
struct PointTree {
    int L, ring, p, scaled;
    const double* root;
    Level lv[kMaxPointLevels];
    GridView gv[kMaxPointLevels];
    const double4* sp;  // sorted ions (FP64 position, charge): near field of outlier points
};

__constant__ float c_fact[kMaxP + 2] = {1.0f,     1.0f,        2.0f,        6.0f,         24.0f,
                                        120.0f,   720.0f,      5040.0f,     40320.0f,     362880.0f,
                                        3628800.0f, 39916800.0f, 479001600.0f, 6227020800.0f};

// Field of one box's expansion at u (units of h; sum q d / r^3 convention,
// still to be divided by h^2), from the lane's three degree slices of T in
// LDS: entry e of the lane at S[e * 64] (S = shared base + lane); slice k
// holds degree n with n % 3 == k, position (b + c)(b + c + 1)/2 + c within
// the degree.
__device__ void m2p_box(const float* __restrict__ M, int p, int scaled, float ux, float uy, float uz, float eps2,
                        float* S, int S1, float& gx, float& gy, float& gz) {
    auto T = [&](int deg, int pos) -> float& { return S[((deg % 3) * S1 + pos) * 64]; };
    const float s = fmaf(ux, ux, fmaf(uy, uy, fmaf(uz, uz, eps2)));
    const float is = 1.0f / s;
    T(0, 0) = __builtin_amdgcn_rsqf(s);
    for (int n = 1; n <= p + 1; n++) {
        const float c1 = (float)(2 * n - 1), c2 = (float)(n - 1), cn = -is / (float)n;
        int pos = 0;
        for (int t = 0; t <= n; t++) {
            const int a = n - t;
            for (int c = 0; c <= t; c++, pos++) {
                const int b = t - c;
                float acc = 0.0f;
                if (a >= 1) acc = fmaf(ux, T(n - 1, t * (t + 1) / 2 + c), acc);
                if (b >= 1) acc = fmaf(uy, T(n - 1, (t - 1) * t / 2 + c), acc);
                if (c >= 1) acc = fmaf(uz, T(n - 1, (t - 1) * t / 2 + c - 1), acc);
                float acc2 = 0.0f;
                if (n >= 2) {
                    if (a >= 2) acc2 += T(n - 2, t * (t + 1) / 2 + c);
                    if (b >= 2) acc2 += T(n - 2, (t - 2) * (t - 1) / 2 + c);
                    if (c >= 2) acc2 += T(n - 2, (t - 2) * (t - 1) / 2 + c - 2);
                }
                T(n, pos) = cn * fmaf(c1, acc, c2 * acc2);
            }
        }
        // terms of the coefficients of degree n - 1 (their T_{a+e_k} have degree n)
        const int m = n - 1;
        const float sg = (m & 1) ? -1.0f : 1.0f;
        int q = 0;
        for (int t = 0; t <= m; t++) {
            const int a0 = m - t;
            for (int c = 0; c <= t; c++, q++) {
                const int a1 = t - c, a2 = c;
                const float Ma = M[dstart(m) + q];
                const float w = scaled ? c_fact[a0] * c_fact[a1] * c_fact[a2] : sg;
                const float wm = w * Ma;
                gx = fmaf(wm * (float)(a0 + 1), T(n, t * (t + 1) / 2 + c), gx);
                gy = fmaf(wm * (float)(a1 + 1), T(n, (t + 1) * (t + 2) / 2 + c), gy);
                gz = fmaf(wm * (float)(a2 + 1), T(n, (t + 1) * (t + 2) / 2 + c + 1), gz);
            }
        }
    }
}

// One wave per point (an ion, by its index in the caller's arrays): adds
// K q / m times the far field of the tree at the point (M2P, levels 0..L)
// and, with near_p2p, the near field of the leaf boxes within `ring` cells of
// its leaf cell (outliers, which are not in the tree).  Points with status
// other than 1 and waves past *npts return at once.
__global__ __launch_bounds__(64) void k_fmm_m2p(const PointTree P, const int* __restrict__ pts,
                                                const int* __restrict__ npts, int near_p2p,
                                                const double* __restrict__ X, const double* __restrict__ Y,
                                                const double* __restrict__ Z, const double* __restrict__ Q,
                                                const int* __restrict__ status, const int* __restrict__ species,
                                                double K, double* __restrict__ ax, double* __restrict__ ay,
                                                double* __restrict__ az) {
    extern __shared__ float sm[];
    __shared__ int list[kM2PList];
    __shared__ int nlist;
    const int lane = threadIdx.x;
    if ((int)blockIdx.x >= *npts) return;
    const int i = pts[blockIdx.x];
    if (status && status[i] != 1) return;  // uniform
    const double x = X[i], y = Y[i], z = Z[i];
    const int S1 = (P.p + 2) * (P.p + 3) / 2;
    float* S = sm + lane;
    const int ring = P.ring, nside = 2 * ring + 1, ncand = nside * nside * nside * 8;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int l = 0; l <= P.L; l++) {
        const double w = P.root[3] / (double)(1 << l);
        const long long cx = (long long)floor((x - P.root[0]) / w), cy = (long long)floor((y - P.root[1]) / w),
                        cz = (long long)floor((z - P.root[2]) / w);
        const long long top = (1LL << l) - 1;
        if (lane == 0) nlist = 0;
        __syncthreads();
        if (l == 0) {
            const bool near = llabs(cx) <= ring && llabs(cy) <= ring && llabs(cz) <= ring;
            if (lane == 0 && (!near || l == P.L)) {
                list[0] = near ? -1 : 0;  // -1: near root leaf (P2P), else far root (M2P)
                nlist = 1;
            }
        } else {
            const long long qx = cx >> 1, qy = cy >> 1, qz = cz >> 1;  // the point's cell one level up (floor)
            for (int t0 = 0; t0 < ncand; t0 += 64) {
                const int t = t0 + lane;
                int entry = 0;  // 0: none; b + 1: far box b; -(b + 1): near leaf box b
                if (t < ncand) {
                    const int po = t >> 3, oc = t & 7;
                    const long long px = qx + po % nside - ring, py = qy + (po / nside) % nside - ring,
                                    pz = qz + po / (nside * nside) - ring;
                    const long long ptop = (1LL << (l - 1)) - 1;
                    if (px >= 0 && py >= 0 && pz >= 0 && px <= ptop && py <= ptop && pz <= ptop) {
                        const long long bx = 2 * px + (oc & 1), by = 2 * py + ((oc >> 1) & 1),
                                        bz = 2 * pz + ((oc >> 2) & 1);
                        const bool near = llabs(bx - cx) <= ring && llabs(by - cy) <= ring && llabs(bz - cz) <= ring;
                        if ((!near || (l == P.L && near_p2p)) && bx <= top && by <= top && bz <= top) {
                            const int b = find_cell(P.lv[l], P.gv[l], (int)bx, (int)by, (int)bz);
                            if (b >= 0) entry = near ? -(b + 1) : b + 1;
                        }
                    }
                }
                const unsigned long long mk = __ballot(entry != 0);
                if (entry != 0) list[nlist + __popcll(mk & ((1ull << lane) - 1ull))] = entry;
                __syncthreads();
                if (lane == 0) nlist += __popcll(mk);
                __syncthreads();
            }
        }
        __syncthreads();
        const int nl = nlist;
        const double h = 0.5 * w;
        const float eps2 = (float)(kSoftening / (h * h));
        for (int k = lane; k < nl; k += 64) {
            const int e = list[k];
            const int b = e > 0 ? e - 1 : (e == 0 ? 0 : -e - 1);
            if (l == 0 && e == -1) continue;  // root is the near leaf: handled below
            if (e > 0 || (l == 0 && e == 0)) {
                double c[3], hb;
                box_center(P.root, l, P.lv[l].key[b], c, hb);
                float gx = 0.0f, gy = 0.0f, gz = 0.0f;
                m2p_box(P.lv[l].M + (size_t)b * P.lv[l].ld, P.p, P.scaled, (float)((x - c[0]) / hb),
                        (float)((y - c[1]) / hb), (float)((z - c[2]) / hb), eps2, S, S1, gx, gy, gz);
                const double ih2 = 1.0 / (hb * hb);
                ex -= (double)gx * ih2;
                ey -= (double)gy * ih2;
                ez -= (double)gz * ih2;
            } else {
                for (int j = P.lv[l].start[b]; j < P.lv[l].end[b]; j++) {
                    const double4 sj = P.sp[j];
                    const double dx = x - sj.x, dy = y - sj.y, dz = z - sj.z;
                    const double r2 = dx * dx + dy * dy + dz * dz + kSoftening;
                    const double inv = rsqrt_f64(r2);
                    const double wq = sj.w * inv * inv * inv;
                    ex += dx * wq;
                    ey += dy * wq;
                    ez += dz * wq;
                }
            }
        }
        if (l == 0 && nl == 1 && list[0] == 0) break;  // far from the whole tree: the root's expansion is all
        if (l == 0 && nl == 1 && list[0] == -1 && near_p2p) {
            // single-level tree (L = 0) with the point near the root: direct sum over the root's ions
            for (int j = P.lv[0].start[0] + lane; j < P.lv[0].end[0]; j += 64) {
                const double4 sj = P.sp[j];
                const double dx = x - sj.x, dy = y - sj.y, dz = z - sj.z;
                const double inv = rsqrt_f64(dx * dx + dy * dy + dz * dz + kSoftening);
                const double wq = sj.w * inv * inv * inv;
                ex += dx * wq;
                ey += dy * wq;
                ez += dz * wq;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    double(*red)[64] = reinterpret_cast<double(*)[64]>(sm);  // the T slices are free now
    red[0][lane] = ex;
    red[1][lane] = ey;
    red[2][lane] = ez;
    __syncthreads();
    if (lane == 0) {
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int k = 0; k < 64; k++) {
            sx += red[0][k];
            sy += red[1][k];
            sz += red[2][k];
        }
        const double kc = K * Q[i] / species_mass(species[i]);
        ax[i] += sx * kc;
        ay[i] += sy * kc;
        az[i] += sz * kc;
    }
}

// The list of the sorted target ions by their original index (treecode:
// every target gets the M2P far field).
__global__ void k_fmm_target_list(int na, const int* __restrict__ tfl, const int* __restrict__ tpre,
                                  const int* __restrict__ perm, int* __restrict__ list) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < na && tfl[i]) list[tpre[i]] = perm[i];
}

__global__ void k_fmm_fill(double* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

// Symmetry images as source-only ions: block 0 holds the ions themselves
// (their own status), block k >= 1 the k-th non-identity image (status 2 for
// ions that are sources, 0 otherwise).  img[k] = (c, s, 0, 0) for a rotation
// about z or (ox, oy, oz, 1) for a translation; img[0] is the identity.
__global__ void k_fmm_images(int n, int nblk, const double4* __restrict__ img,
                             const double* __restrict__ X, const double* __restrict__ Y,
                             const double* __restrict__ Z, const double* __restrict__ Q,
                             const int* __restrict__ st, const int* __restrict__ sp,
                             double* __restrict__ X2, double* __restrict__ Y2, double* __restrict__ Z2,
                             double* __restrict__ Q2, int* __restrict__ st2, int* __restrict__ sp2) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * nblk) return;
    const int k = (int)(t / n), i = (int)(t % n);
    const double4 m = img[k];
    const double x = X[i], y = Y[i], z = Z[i];
    if (m.w != 0.0) {
        X2[t] = x + m.x;
        Y2[t] = y + m.y;
        Z2[t] = z + m.z;
    } else {
        X2[t] = m.x * x - m.y * y;
        Y2[t] = m.y * x + m.x * y;
        Z2[t] = z;
    }
    Q2[t] = Q[i];
    sp2[t] = sp[i];
    const int s = st[i];
    st2[t] = k == 0 ? s : ((s == 1 || s == 2) ? 2 : 0);
}

// ------------------------------------------------------------- outliers ----
// DomainBoundsFunction::isOutlier (domain_bounds_functions.cpp:16-93): per
// axis, the median of the valid ions (status 1 or 2) and their median
// absolute deviation times outlierFactor (1.4826, as a float); an ion is an
// outlier when its deviation is at least maxDeviation (3) such units on some
// axis.  The reference's selection drops values on ties and takes the
// deviation median over the whole range, valid or not (outlier_detection.h:
// 61-82, domain_bounds_functions.cpp:62); here both medians are those of the
// valid ions, read from sorted keys, in FP64.
constexpr double kOutlierFactor = (double)1.482602218505602f;
constexpr double kMaxDeviation = 3.0;

// Sort keys of one axis: the coordinate (med == nullptr) or its deviation
// from the median; invalid ions sort last (+inf).  Counts the valid ions.
__global__ void k_outl_keys(int n, const double* __restrict__ v, const int* __restrict__ st,
                            const double* __restrict__ med, double* __restrict__ key, int* __restrict__ nvalid) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = st[i];
    const bool valid = s == 1 || s == 2;
    key[i] = valid ? (med ? fabs(v[i] - *med) : v[i]) : INFINITY;
    if (nvalid && valid) atomicAdd(nvalid, 1);
}

// out = scale x the median of the nv smallest sorted keys (the mean of the two
// middle ones for even nv, findTrueMedian, outlier_detection.h:118-133).
__global__ void k_outl_median(const double* __restrict__ sorted, const int* __restrict__ nvalid, double scale,
                              double* __restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int m = *nvalid;
    *out = m > 0 ? scale * (0.5 * (sorted[(m - 1) / 2] + sorted[m / 2])) : 0.0;
}

// stat = {median x, y, z, MAD x, y, z} (MAD already times outlierFactor).
__global__ void k_outl_flags(int n, const double* __restrict__ X, const double* __restrict__ Y,
                             const double* __restrict__ Z, const int* __restrict__ st,
                             const double* __restrict__ stat, int* __restrict__ flag, int* __restrict__ pflag,
                             int* __restrict__ core) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = st[i];
    const bool valid = s == 1 || s == 2;
    const double v[3] = {X[i], Y[i], Z[i]};
    bool out = false;
    for (int c = 0; c < 3; c++) out |= fabs(v[c] - stat[c]) / stat[3 + c] >= kMaxDeviation;
    out = out && valid;
    flag[i] = out;
    pflag[i] = out && s == 1;  // outliers that are targets
    core[i] = out ? 0 : s;
}

// Targets (status 1) += K q/m x the direct FP64 sum over the outlier sources
// (outlier_p2p, fmm_cpukernel.cpp:2230-2280: every target against the
// outliers), self excluded; the outliers staged in LDS 256 at a time.
__global__ __launch_bounds__(256) void k_outl_sources(int n, const double* __restrict__ X,
                                                      const double* __restrict__ Y, const double* __restrict__ Z,
                                                      const double* __restrict__ Q, const int* __restrict__ st,
                                                      const int* __restrict__ species, const int* __restrict__ olist,
                                                      int nout, double K, double* __restrict__ ax,
                                                      double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double4 src[256];
    __shared__ int sid[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool tgt = i < n && st[i] == 1;
    const double x = tgt ? X[i] : 0.0, y = tgt ? Y[i] : 0.0, z = tgt ? Z[i] : 0.0;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j0 = 0; j0 < nout; j0 += 256) {
        __syncthreads();
        if (j0 + (int)threadIdx.x < nout) {
            const int j = olist[j0 + threadIdx.x];
            src[threadIdx.x] = make_double4(X[j], Y[j], Z[j], Q[j]);
            sid[threadIdx.x] = j;
        }
        __syncthreads();
        const int cnt = min(256, nout - j0);
        if (tgt)
            for (int k = 0; k < cnt; k++) {
                if (sid[k] == i) continue;
                const double4 s = src[k];
                const double dx = x - s.x, dy = y - s.y, dz = z - s.z;
                const double inv = rsqrt_f64(dx * dx + dy * dy + dz * dz + kSoftening);
                const double w = s.w * inv * inv * inv;
                sx += dx * w;
                sy += dy * w;
                sz += dz * w;
            }
    }
    if (tgt && nout > 0) {
        const double kc = K * Q[i] / species_mass(species[i]);
        ax[i] += sx * kc;
        ay[i] += sy * kc;
        az[i] += sz * kc;
    }
}

// Folds the measurements of the last FMM call into the handle: its events
// and counters are complete once the stream has passed them (the next call's
// host readback, or an explicit wait here).
static int fmm_fold(shp_handle h) {
    if (!h->fmm_pending) return SHP_OK;
    SHP_HIP_CHECK(hipEventSynchronize(h->fev[5]));
    double v[kFmmStats] = {};
    for (int k = 0; k < 5; k++) {
        float ms = 0.0f;
        SHP_HIP_CHECK(hipEventElapsedTime(&ms, h->fev[k], h->fev[k + 1]));
        v[k] = ms;
    }
    unsigned long long c[5];
    SHP_HIP_CHECK(hipMemcpy(c, h->fmm_stats, sizeof(c), hipMemcpyDeviceToHost));
    v[5] = (double)c[0];
    v[6] = (double)c[1];
    v[7] = h->fmm_leaf_level;
    if (h->fmm_unrolled) {
        for (int k = 0; k < 3; k++) {
            v[8] += (double)c[2 + k] * h->fmm_shell_madds[k];
            v[10] += (double)c[2 + k] * h->fmm_shell_madds0[k];
        }
    } else {
        v[8] = v[10] = (double)c[0] * h->fmm_shell_madds[0];
    }
    float ms0 = 0.0f;
    SHP_HIP_CHECK(hipEventElapsedTime(&ms0, h->fev[6], h->fev[7]));
    v[9] = ms0;
    for (int k = 0; k < kFmmStats; k++) {
        h->fmm_last[k] = v[k];
        h->fmm_acc[k] += v[k];
    }
    h->fmm_acc_n++;
    h->fmm_pending = false;
    h->fmm_have_last = true;
    return SHP_OK;
}

}  // namespace shp

using namespace shp;

extern "C" {

// Points outside the tree that get the tree's field through M2P plus the near
// field of their leaf cell (the outliers of the *Outliers modes): original
// indices, their number in device memory, and an upper bound of it.
struct FmmPoints {
    const int* idx;
    const int* count;
    int max;
};

static int fmm_eval(shp_handle h, const shp_particles* pt, int order, int leaf_level, int leaf_size, int ring,
                    bool full, double* ax, double* ay, double* az, const FmmPoints* pts = nullptr) {
    SHP_REQUIRE(h && pt && ax && ay && az, "shp_fmm_dev: null argument");
    SHP_REQUIRE(order >= 1 && order <= kMaxP, "shp_fmm_dev: order %d outside [1, %d]", order, kMaxP);
    SHP_REQUIRE(ring >= 1 && ring <= 2, "shp_fmm_dev: ring %d outside [1, 2]", ring);
    SHP_REQUIRE(leaf_level >= 0 && leaf_level <= 16, "shp_fmm_dev: leaf level %d outside [0, 16]",
                leaf_level);
    SHP_REQUIRE(leaf_level > 0 || leaf_size >= 1, "shp_fmm_dev: leaf_size %d must be >= 1",
                leaf_size);
    const int n = pt->n;
    if (n <= 0) return SHP_OK;
    SHP_REQUIRE(pt->x && pt->y && pt->z && pt->q && pt->status && pt->species,
                "shp_fmm_dev: null particle array");
    // treecode: the far field of every target by M2P, no local expansions
    const bool treecode = h->fmm_type == SHP_FMM_TREECODE || h->fmm_type == SHP_FMM_TREECODE_OUTLIERS;
    if ((treecode || pts) && order > kM2PMaxP) {
        set_error("shp_fmm_dev: treecode and outlier modes evaluate multipoles up to order %d (order %d)",
                  kM2PMaxP, order);
        return SHP_ERR_UNSUPPORTED;
    }
    SHP_HIP_CHECK(hipSetDevice(h->device));
    FmmDevTables* tb;
    int rc = get_tables(h, order, &tb);
    if (rc) return rc;
    int L = leaf_level;
    const int nc = tb->nc;
    hipStream_t st = h->stream;
    // bounding box and root
    const int nbp = std::min(fdiv_up(n, kFmmThreads), 1024);
    SHP_WS(bpart, double, h, WS_BBOX, nbp * 6);
    SHP_WS(root, double, h, WS_BOX, 16);
    hipLaunchKernelGGL(k_fmm_bbox_partial, dim3(nbp), dim3(kFmmThreads), 0, st, pt->x, pt->y, pt->z,
                       pt->status, n, bpart);
    hipLaunchKernelGGL(k_fmm_root, dim3(1), dim3(64), 0, st, bpart, nbp, root);
    // keys and sort
    SHP_WS(keys, uint64_t, h, WS_FMM_KEYS, n);
    SHP_WS(keys2, uint64_t, h, WS_FMM_KEYS2, n);
    SHP_WS(vals, int, h, WS_TGT, n);
    SHP_WS(perm, int, h, WS_VALS2, n);
    hipLaunchKernelGGL(k_fmm_keys, dim3(fdiv_up(n, 256)), dim3(256), 0, st, pt->x, pt->y, pt->z,
                       pt->status, n, root, keys, vals);
    size_t tmp = 0;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, keys, keys2, vals, perm, n, 0, 64, st));
    SHP_WS(cub, char, h, WS_CUB, tmp);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub, tmp, keys, keys2, vals, perm, n, 0, 64, st));
    // box counts per level (the host needs them for launch sizes)
    SHP_WS(cnt, int, h, WS_COUNTS, 32);
    hipLaunchKernelGGL(k_fmm_ifill, dim3(1), dim3(32), 0, st, cnt, 32, 0);
    hipLaunchKernelGGL(k_fmm_level_hist, dim3(std::min(fdiv_up(n, 256), 512)), dim3(256), 0, st, keys2, n, cnt);
    int hc[32];
    double hroot[10];
    SHP_HIP_CHECK(hipMemcpyAsync(hc, cnt, sizeof(int) * 32, hipMemcpyDeviceToHost, st));
    SHP_HIP_CHECK(hipMemcpyAsync(hroot, root, sizeof(double) * 10, hipMemcpyDeviceToHost, st));
    SHP_HIP_CHECK(hipStreamSynchronize(st));
    {
        // the previous call's events and counters are complete: fold them
        // before this call resets the counters
        const int frc = fmm_fold(h);
        if (frc != SHP_OK) return frc;
    }
    const int na = hc[0];
    std::vector<int> nbox_all(kMaxLevel + 1);
    for (int l = 0, acc = 0; l <= kMaxLevel; l++) {
        acc += hc[1 + l];
        nbox_all[l] = acc;
    }
    // smallest level whose mean leaf occupancy is at most leaf_size
    int Lrule = 1;
    while (Lrule < 16 && (double)na / std::max(1, nbox_all[Lrule]) > (double)leaf_size) Lrule++;
    if (na == 0) {
        last_leaf_level(h) = L > 0 ? L : Lrule;
        hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, ax, n);
        hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, ay, n);
        hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, az, n);
        return SHP_OK;
    }
    int p3 = 0, p4 = 0, p5 = 0;
    m2l_orders(order, ring, full, p3, p4, p5);
    // Dense cell -> box grids per level over the occupied cells (unrolled M2L,
    // near-field lookups, leaf-level choice); very large grids fall back to the
    // table-driven M2L and key searches.
    std::vector<int> glo, gdim;
    auto grid_cells = [&](int Lmax) {
        glo.assign(3 * (Lmax + 1), 0);
        gdim.assign(3 * (Lmax + 1), 1);
        size_t cells_tot = 0;
        for (int l = 0; l <= Lmax; l++) {
            const double sc = std::ldexp(1.0, l) / hroot[3];
            size_t cells = 1;
            for (int c = 0; c < 3; c++) {
                const int top = (1 << l) - 1;
                int lo = (int)std::floor((hroot[4 + c] - hroot[c]) * sc) - 1;
                int hi = (int)std::floor((hroot[7 + c] - hroot[c]) * sc) + 1;
                lo = std::max(0, std::min(lo, top));
                hi = std::max(lo, std::min(hi, top));
                glo[3 * l + c] = lo;
                gdim[3 * l + c] = hi - lo + 1;
                cells *= (size_t)(hi - lo + 1);
            }
            cells_tot += cells;
        }
        return cells_tot;
    };
    constexpr size_t kGridCells = (size_t)1 << 27;
    // With leaf_level 0 and the unrolled M2L, the levels are built one deeper
    // than the occupancy rule and the leaf level is chosen by estimated cost.
    int Lb = L > 0 ? L : Lrule;
    const bool choose = L <= 0 && p3 > 0 && Lrule >= 2 && Lrule < 16 && grid_cells(Lrule + 1) <= kGridCells;
    if (choose) Lb = Lrule + 1;
    const size_t gtot = grid_cells(Lb);
    const bool unrolled = p3 > 0 && gtot <= kGridCells;
    std::vector<int> nbox(Lb + 1);
    for (int l = 0; l <= Lb; l++) nbox[l] = nbox_all[l];
    // sorted particle records
    SHP_WS(sp, double4, h, WS_SRC, na);
    hipLaunchKernelGGL(k_fmm_gather, dim3(fdiv_up(na, 256)), dim3(256), 0, st, perm, na, pt->x, pt->y,
                       pt->z, pt->q, sp);
    // target prefix sums over the sorted ions (boxes without targets skip the
    // M2L, L2L and leaf work)
    SHP_WS(tfl, int, h, WS_FMM_TPRE, 2 * ((size_t)na + 1));
    int* const tpre = tfl + na + 1;
    {
        hipLaunchKernelGGL(k_fmm_tflag, dim3(fdiv_up(na + 1, 256)), dim3(256), 0, st, perm, na, pt->status, tfl);
        size_t tsz = 0;
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tsz, tfl, tpre, na + 1, st));
        SHP_WS(cubp, char, h, WS_CUB, tsz);
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cubp, tsz, tfl, tpre, na + 1, st));
    }
    // level arrays in one workspace: key (u64), start, end, cfirst per box
    size_t totb = 0;
    for (int l = 0; l <= Lb; l++) totb += nbox[l];
    SHP_WS(bkey, uint64_t, h, WS_FMM_BKEY, totb);
    SHP_WS(bint, int, h, WS_FMM_BINT, 3 * totb);
    SHP_WS(flags, char, h, WS_FLAGS, na);
    std::vector<Level> lv(Lb + 1);
    {
        size_t off = 0;
        for (int l = 0; l <= Lb; l++) {
            Level& v = lv[l];
            v.nb = nbox[l];
            v.key = bkey + off;
            v.start = bint + off;
            v.end = bint + totb + off;
            v.cfirst = l < Lb ? bint + 2 * totb + off : nullptr;
            v.M = nullptr;
            v.Lc = nullptr;
            v.tpre = tpre;
            off += nbox[l];
        }
    }
    // The leaf-most level from the sorted particle keys; each coarser level
    // from the boxes of the level below (a select over boxes, not particles),
    // which also gives the first-child index directly.
    int lup = -1;  // the top levels lup .. 0 are built by one single-workgroup launch
    for (int l = Lb; l >= 0; l--) {
        hipcub::CountingInputIterator<int> it(0);
        size_t t2 = 0;
        if (l < Lb && nbox[l + 1] <= kUpperMaxChildren) {
            lup = l;
            break;
        }
        if (l == Lb) {
            const int sh = 3 * (kMaxLevel - l);
            hipLaunchKernelGGL(k_fmm_flags, dim3(fdiv_up(na, 256)), dim3(256), 0, st, keys2, na, sh, flags);
            SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, t2, it, flags, (int*)lv[l].start, cnt + 24,
                                                        na, st));
            SHP_WS(cub2, char, h, WS_CUB, t2);
            SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub2, t2, it, flags, (int*)lv[l].start, cnt + 24,
                                                        na, st));
            hipLaunchKernelGGL(k_fmm_box_info, dim3(fdiv_up(nbox[l], 256)), dim3(256), 0, st, keys2,
                               lv[l].start, nbox[l], na, sh, (uint64_t*)lv[l].key, (int*)lv[l].end);
        } else {
            const int ncb = nbox[l + 1];
            hipLaunchKernelGGL(k_fmm_parent_flags, dim3(fdiv_up(ncb, 256)), dim3(256), 0, st, lv[l + 1].key, ncb,
                               flags);
            SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, t2, it, flags, (int*)lv[l].cfirst, cnt + 24,
                                                        ncb, st));
            SHP_WS(cub2, char, h, WS_CUB, t2);
            SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub2, t2, it, flags, (int*)lv[l].cfirst, cnt + 24,
                                                        ncb, st));
            hipLaunchKernelGGL(k_fmm_parent_info, dim3(fdiv_up(nbox[l], 256)), dim3(256), 0, st, lv[l].cfirst,
                               nbox[l], na, lv[l + 1].key, lv[l + 1].start, (uint64_t*)lv[l].key, (int*)lv[l].start,
                               (int*)lv[l].end);
        }
    }
    if (lup >= 0) {
        UpperLevels U{};
        U.lhi = lup;
        U.na = na;
        for (int l = 0; l <= lup + 1; l++) {
            U.key[l] = (uint64_t*)lv[l].key;
            U.start[l] = (int*)lv[l].start;
            U.end[l] = (int*)lv[l].end;
            U.cfirst[l] = (int*)lv[l].cfirst;
            U.nb[l] = nbox[l];
        }
        hipLaunchKernelGGL(k_fmm_upper_levels, dim3(1), dim3(1024), 0, st, U);
    }
    SHP_HIP_CHECK(hipGetLastError());
    std::vector<int*> gl(Lb + 1, nullptr);
    auto gview = [&](int l) {
        GridView g{};
        if (unrolled) {
            g.g = gl[l];
            for (int c = 0; c < 3; c++) {
                g.lo[c] = glo[3 * l + c];
                g.dim[c] = gdim[3 * l + c];
            }
        }
        return g;
    };
    if (unrolled) {
        SHP_WS(grid, int, h, WS_FMM_GRID, gtot);
        hipLaunchKernelGGL(k_fmm_ifill, dim3(fdiv_up((long)gtot, 256)), dim3(256), 0, st, grid, (int)gtot, -1);
        SHP_REQUIRE(Lb + 1 <= kGridLevels, "shp_fmm_dev: %d levels", Lb + 1);
        GridScatter S{};
        S.key = lv[0].key;
        S.nl = Lb + 1;
        size_t goff = 0;
        int boff = 0;
        for (int l = 0; l <= Lb; l++) {
            gl[l] = grid + goff;
            S.boff[l] = boff;
            S.goff[l] = (long long)goff;
            for (int c = 0; c < 3; c++) {
                S.lo[l][c] = glo[3 * l + c];
                S.dim[l][c] = gdim[3 * l + c];
            }
            SHP_REQUIRE(lv[l].key == S.key + boff, "shp_fmm_dev: level key arrays not contiguous");
            boff += nbox[l];
            goff += (size_t)gdim[3 * l] * gdim[3 * l + 1] * gdim[3 * l + 2];
        }
        S.boff[Lb + 1] = boff;
        hipLaunchKernelGGL(k_fmm_grid_scatter, dim3(fdiv_up(boff, 256)), dim3(256), 0, st, S, grid);
    }
    if (choose) {
        // P2P pairs and M2L far-band box pairs of the candidate leaf levels
        // Lrule - 2 .. Lrule + 1, counted on every stride-th box (at most
        // 4096 boxes per level; the sums are scaled to the whole level).
        const int L0 = std::max(2, Lrule - 2);
        SHP_WS(lcost, unsigned long long, h, WS_FMM_LCOST, 2 * (Lb + 1));
        hipLaunchKernelGGL(k_fmm_u64fill, dim3(1), dim3(64), 0, st, lcost, 2 * (Lb + 1));
        std::vector<int> stride(Lb + 1, 1);
        LevelCostArgs A{};
        int ns_max = 1;
        for (int l = L0; l <= Lb; l++) {
            const int k = l - L0;
            stride[l] = std::max(1, fdiv_up(nbox[l], 1 << 12));
            ns_max = std::max(ns_max, fdiv_up(nbox[l], stride[l]));
            A.lv[k] = lv[l];
            A.par[k] = lv[l - 1];
            A.g[k] = gview(l);
            A.gp[k] = gview(l - 1);
            A.stride[k] = stride[l];
            A.out[k] = lcost + 2 * l;
        }
        SHP_REQUIRE(Lb - L0 + 1 <= kLevelCostMax, "shp_fmm_dev: %d candidate levels", Lb - L0 + 1);
        hipLaunchKernelGGL(k_fmm_level_cost, dim3(std::min(kLevelCostBlocks, fdiv_up(ns_max, 4)), Lb - L0 + 1),
                           dim3(256), 0, st, A, ring);
        unsigned long long hcost[2 * (kMaxLevel + 1)];
        SHP_HIP_CHECK(hipMemcpyAsync(hcost, lcost, sizeof(unsigned long long) * 2 * (Lb + 1),
                                     hipMemcpyDeviceToHost, st));
        SHP_HIP_CHECK(hipStreamSynchronize(st));
        double best = 0.0, m2l = 0.0;
        h->fmm_est_n = 0;
        for (int l = L0; l <= Lb; l++) {
            const double sc = (double)nbox[l] / fdiv_up(nbox[l], stride[l]);
            m2l += fmm_m2l_cost_ms(nc, nbox[l], sc * (double)hcost[2 * l + 1]);
            const double c = m2l + fmm_leaf_cost_ms(nbox[l], sc * (double)hcost[2 * l]);
            double* row = h->fmm_est + 4 * h->fmm_est_n++;
            row[0] = l;
            row[1] = sc * (double)hcost[2 * l];
            row[2] = sc * (double)hcost[2 * l + 1];
            row[3] = c;
            if (l == L0 || c < best) {
                best = c;
                L = l;
            }
        }
    } else {
        h->fmm_est_n = 0;
        if (L <= 0) L = Lrule;
    }
    last_leaf_level(h) = L;
    nbox.resize(L + 1);
    lv.resize(L + 1);
    lv[L].cfirst = nullptr;  // leaf level
    // expansions M, L per box of levels 0..L
    size_t tot = 0;
    for (int l = 0; l <= L; l++) tot += nbox[l];
    const int ld = (nc + 31) & ~31;  // expansion rows padded to whole 128-byte lines
    SHP_WS(bexp, float, h, WS_FMM_EXP, 2 * tot * (size_t)ld);
    {
        size_t off = 0;
        for (int l = 0; l <= L; l++) {
            lv[l].M = bexp + off * ld;
            lv[l].Lc = bexp + (tot + off) * ld;
            lv[l].ld = ld;
            off += nbox[l];
        }
    }
    SHP_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_fmm_zero, dim3(fdiv_up(tot * ld, 256)), dim3(256), 0, st, bexp + tot * ld,
                       tot * ld);
    hipLaunchKernelGGL(k_fmm_u64fill, dim3(1), dim3(64), 0, st, h->fmm_stats, 8);
    hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, ax, n);
    hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, ay, n);
    hipLaunchKernelGGL(k_fmm_fill, dim3(fdiv_up(n, 256)), dim3(256), 0, st, az, n);
    // leaf work units (at most nbox + na / kLeafUnit of them), longest first
    const int umax = nbox[L] + na / kLeafUnit + 1;
    SHP_WS(ucnt, int, h, WS_FMM_UCNT, 3 * (size_t)nbox[L] + 1);
    SHP_WS(units, int2, h, WS_FMM_UNITS, (size_t)umax);
    SHP_WS(usort, int, h, WS_FMM_USORT, 4 * (size_t)umax);
    int* const nunits = ucnt + 3 * nbox[L];
    int* const uorder = usort + 3 * (size_t)umax;
    {
        const int nbl = nbox[L];
        int* uoffs = ucnt + nbl;
        int* nsrcb = ucnt + 2 * nbl;
        hipLaunchKernelGGL(k_fmm_unit_count, dim3(fdiv_up(nbl, 4)), dim3(256), 0, st, lv[L], gview(L), L, ring,
                           ucnt, nsrcb);
        size_t tsz = 0;
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tsz, ucnt, uoffs, nbl, st));
        SHP_WS(cubu, char, h, WS_CUB, tsz);
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cubu, tsz, ucnt, uoffs, nbl, st));
        hipLaunchKernelGGL(k_fmm_unit_fill, dim3(fdiv_up(nbl, 256)), dim3(256), 0, st, ucnt, uoffs, nbl, units,
                           nunits);
        unsigned* ukey = (unsigned*)usort;
        int* uval = usort + 2 * (size_t)umax;
        hipLaunchKernelGGL(k_fmm_unit_keys, dim3(fdiv_up(umax, 256)), dim3(256), 0, st, units, nunits, umax, lv[L],
                           nsrcb, ukey, uval);
        size_t ssz = 0;
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, ssz, ukey, ukey + umax, uval, uorder,
                                                                   umax, 0, 32, st));
        SHP_WS(cubs, char, h, WS_CUB, ssz);
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairsDescending(cubs, ssz, ukey, ukey + umax, uval, uorder,
                                                                   umax, 0, 32, st));
    }
    SHP_WS(spf, float4, h, WS_FMM_SPF, na);
    hipLaunchKernelGGL(k_fmm_leaf_rel, dim3(fdiv_up(na, 256)), dim3(256), 0, st, na, L, keys2, root, sp, spf);
    // upward pass
    SHP_HIP_CHECK(hipEventRecord(h->ev0, st));
    SHP_HIP_CHECK(hipEventRecord(h->fev[0], st));
    hipLaunchKernelGGL(k_fmm_p2m, dim3(nbox[L]), dim3(kFmmThreads), 0, st, lv[L], L, root, sp, nc,
                       order, tb->mi);
    for (int l = L - 1; l >= 0; l--)
        hipLaunchKernelGGL(k_fmm_m2m, dim3(nbox[l]), dim3(kShiftLanes), 8 * nc * sizeof(float), st,
                           lv[l], lv[l + 1], nc, tb->m2m_term, tb->m2m_nt);
    SHP_HIP_CHECK(hipEventRecord(h->fev[1], st));
    // translation tables for offsets within the far band
    const int W = 2 * ring + 1;
    const int nside = 2 * W + 1, noff = nside * nside * nside;
    SHP_WS(ttab, float, h, WS_FMM_T, (size_t)noff * (L + 1) * nc);
    if (!treecode)
        hipLaunchKernelGGL(k_fmm_ttable, dim3(noff * (L + 1)), dim3(64), 0, st, order, nc,
                           tb->mi, W, 0, L, root, kSoftening, ttab);
    // Unrolled M2L (ring 2, orders with compiled pair lists) needs a dense box
    // grid per level over the occupied cells; very large grids use the
    // table-driven kernel instead.
    const GridView leafgrid = gview(L);  // leaf-level cell grid for the near-field lookups
    SHP_HIP_CHECK(hipEventRecord(h->fev[2], st));
    if (treecode) {
        // no M2L: the leaf kernel's L2P adds zero, the M2P pass below the far field
        SHP_HIP_CHECK(hipEventRecord(h->fev[6], st));
        SHP_HIP_CHECK(hipEventRecord(h->fev[7], st));
    } else if (unrolled) {
        const dim3 rowb(64 * fdiv_up(nc, 64));  // one lane per coefficient (nc <= 455)
        hipLaunchKernelGGL(k_fmm_scale, dim3((unsigned)tot), rowb, 0, st, bexp, tb->mscale, nc, ld);
        hipLaunchKernelGGL(k_fmm_scale, dim3((unsigned)(noff * (L + 1))), rowb, 0, st, ttab, tb->dscale, nc, nc);
        if (L >= 2) {
            M2LPlan pl{};
            pl.nc = nc;
            pl.ld = ld;
            size_t gbase = 0;
            for (int l = 0; l < 2; l++) gbase += nbox[l];
            const size_t first = gbase;
            LevelOffs lofs{};
            int ub = 0;  // tiles of the (level, octant) groups, at most
            for (int l = 2; l <= L; l++) {
                M2LLevel& m = pl.lv[pl.nlev];
                m.l = l;
                m.nb = nbox[l];
                m.gbase = (long)(gbase - first);
                m.key = lv[l].key;
                m.Mhat = lv[l].M;
                m.Lc = lv[l].Lc;
                m.D = ttab + (size_t)l * noff * nc;
                m.grid = gl[l];
                for (int c = 0; c < 3; c++) {
                    m.lo[c] = glo[3 * l + c];
                    m.dim[c] = gdim[3 * l + c];
                }
                lofs.off[pl.nlev] = (int)(gbase - first);
                ub += fdiv_up(nbox[l], 64) + 8;
                pl.nlev++;
                gbase += nbox[l];
            }
            lofs.nl = pl.nlev;
            lofs.off[pl.nlev] = (int)(gbase - first);
            // group the boxes by (level, octant), Morton order within a group
            const int nb2 = (int)(gbase - first);
            const int nseg = 8 * pl.nlev;
            SHP_WS(okey, unsigned short, h, WS_FMM_OKEY, 2 * (size_t)nb2);
            SHP_WS(oval, int, h, WS_FMM_OVAL, 2 * (size_t)nb2);
            SHP_WS(segb, int, h, WS_FMM_SEG, 2 * (size_t)nseg);
            SHP_WS(tiles, int4, h, WS_FMM_TILES, (size_t)ub);
            const int* bst = bint + first;
            const int* ben = bint + totb + first;
            hipLaunchKernelGGL(k_fmm_oct_keys, dim3(fdiv_up(nb2, 256)), dim3(256), 0, st, bkey + first, lofs,
                               bst, ben, tpre, okey, oval);
            size_t tsz = 0;
            SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, tsz, okey, okey + nb2, oval, oval + nb2,
                                                             nb2, 0, 8, st));
            SHP_WS(cubt, char, h, WS_CUB, tsz);
            SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cubt, tsz, okey, okey + nb2, oval, oval + nb2,
                                                             nb2, 0, 8, st));
            hipLaunchKernelGGL(k_fmm_ifill, dim3(fdiv_up(2 * nseg, 256)), dim3(256), 0, st, segb, 2 * nseg, 0);
            hipLaunchKernelGGL(k_fmm_seg_bounds, dim3(fdiv_up(nb2, 256)), dim3(256), 0, st, okey + nb2, nb2,
                               nseg, segb, segb + nseg);
            hipLaunchKernelGGL(k_fmm_tile_table, dim3(1), dim3(64), 0, st, segb, segb + nseg, nseg, ub, tiles);
            pl.tiles = tiles;
            pl.sorted = oval + nb2;
            pl.ntile = ub;
            // Enough work units to give every SIMD several waves; the partial
            // sums of the offset chunks take S x (boxes of levels >= 2) x nc floats.
            const size_t nrow = gbase - first;
            int S = std::max(1, std::min(32, fdiv_up(8 * 4 * h->num_cus, std::max(1, pl.ntile))));
            while (S > 1 && (size_t)S * nrow * ld * sizeof(float) > ((size_t)4 << 30)) S--;
            pl.nchunk = S;
            const int* cb = nullptr;
            SHP_REQUIRE(m2l_bounds(tb, p3, p4, p5, S, &cb) == SHP_OK, "shp_fmm_dev: M2L chunk table");
            pl.nol = tb->olist_n;  // set by m2l_bounds on first use
            pl.s4 = tb->olist_s4;
            pl.s5 = tb->olist_s5;
            float* part = nullptr;
            const size_t stride = nrow * ld;
            if (S > 1) {
                SHP_WS(pw, float, h, WS_FMM_PART, (size_t)S * stride);
                part = pw;
            }
            launch_m2l_tiled(pl, p3, p4, tb, cb, part, stride, h->fmm_stats, st, h->fev[6], h->fev[7]);
            if (S > 1)
                hipLaunchKernelGGL(k_fmm_m2l_reduce, dim3(fdiv_up((long)(stride / 4), 256)), dim3(256), 0, st, part,
                                   stride, S, nc, ld, tb->lscale, lv[2].Lc, stride / 4, bst, ben, tpre);
        }
        if (L < 2) {
            SHP_HIP_CHECK(hipEventRecord(h->fev[6], st));
            SHP_HIP_CHECK(hipEventRecord(h->fev[7], st));
        }
    } else {
        // table-driven M2L; timed as the first pass
        SHP_HIP_CHECK(hipEventRecord(h->fev[6], st));
        for (int l = 1; l <= L; l++)
            hipLaunchKernelGGL(k_fmm_m2l, dim3(nbox[l]), dim3(kFmmThreads), 0, st, lv[l], lv[l - 1], l,
                               ring, nc, tb->ntask, tb->task, tb->pair_alpha, tb->pair_gamma, tb->pair_w,
                               tb->beta_tasks, ttab + (size_t)l * noff * nc, W, h->fmm_stats);
        SHP_HIP_CHECK(hipEventRecord(h->fev[7], st));
    }
    SHP_HIP_CHECK(hipGetLastError());
    SHP_HIP_CHECK(hipEventRecord(h->fev[3], st));
    for (int l = 1; l < L && !treecode; l++)
        hipLaunchKernelGGL(k_fmm_l2l, dim3(nbox[l]), dim3(kShiftLanes), 9 * nc * sizeof(float), st,
                           lv[l], lv[l + 1], nc, tb->mi, tb->l2l_term, tb->l2l_nt);
    SHP_HIP_CHECK(hipEventRecord(h->fev[4], st));
    // whole runs of 8 x kLeafXcdRun blocks (the block -> unit map permutes within a run)
    const int ublocks = fdiv_up(umax, 8 * kLeafXcdRun) * 8 * kLeafXcdRun;
    hipLaunchKernelGGL(k_fmm_leaf, dim3(ublocks), dim3(64), 0, st, lv[L], L, ring, leafgrid, root, sp, perm,
                       pt->status, pt->species, spf, tb->mi, nc, order, k_const_host(), units, uorder, nunits,
                       h->fmm_stats,
                       ax, ay, az);
    SHP_HIP_CHECK(hipGetLastError());
    if (treecode || pts) {
        PointTree P{};
        P.L = L;
        P.ring = ring;
        P.p = order;
        P.scaled = unrolled && !treecode;  // the unrolled M2L scales the multipoles in place
        P.root = root;
        P.sp = sp;
        for (int l = 0; l <= L; l++) {
            P.lv[l] = lv[l];
            P.gv[l] = gview(l);
        }
        const size_t shm = sizeof(float) * 3 * (size_t)((order + 2) * (order + 3) / 2) * 64;
        if (treecode) {
            SHP_WS(tl, int, h, WS_FMM_PTS, (size_t)na);
            hipLaunchKernelGGL(k_fmm_target_list, dim3(fdiv_up(na, 256)), dim3(256), 0, st, na, tfl, tpre, perm, tl);
            hipLaunchKernelGGL(k_fmm_m2p, dim3(na), dim3(64), shm, st, P, tl, tpre + na, 0, pt->x, pt->y, pt->z,
                               pt->q, (const int*)nullptr, pt->species, k_const_host(), ax, ay, az);
        }
        if (pts && pts->max > 0)
            hipLaunchKernelGGL(k_fmm_m2p, dim3(pts->max), dim3(64), shm, st, P, pts->idx, pts->count, 1, pt->x,
                               pt->y, pt->z, pt->q, (const int*)nullptr, pt->species, k_const_host(), ax, ay, az);
        SHP_HIP_CHECK(hipGetLastError());
    }
    SHP_HIP_CHECK(hipEventRecord(h->fev[5], st));
    SHP_HIP_CHECK(hipEventRecord(h->ev1, st));
    h->timed = true;
    h->fmm_pending = true;
    // multiply-adds per M2L box pair for each shell (or the single order), all
    // passes and the first pass (output degrees 0..4)
    auto madds_of = [](int p) { return p <= 0 ? 0.0 : binom_d(p + 6, 6); };
    auto madds0_of = [](int p) { return p <= 0 ? 0.0 : (double)pass_pairs(p, kPassLo[0], kPassHi[0]); };
    h->fmm_unrolled = unrolled;
    h->fmm_leaf_level = L;
    const int ps[3] = {p3, p4, p5};
    for (int k = 0; k < 3; k++) {
        h->fmm_shell_madds[k] = unrolled ? madds_of(ps[k]) : (double)tb->npairs;
        h->fmm_shell_madds0[k] = unrolled ? madds0_of(ps[k]) : (double)tb->npairs;
    }
    return SHP_OK;
}

// Outlier flags of the valid ions (status 1 or 2), the core status (outliers
// set to 0), the outlier list and the list of outlier targets; the number of
// outliers on the host.
struct OutlierSplit {
    int* flag;
    int* core;
    int* olist;   // every outlier (sources), count nout
    int* plist;   // outliers with status 1 (targets), count *pcount
    int* pcount;  // device
    int nout;
};

static int fmm_outliers(shp_handle h, const shp_particles* pt, OutlierSplit* o) {
    const int n = pt->n;
    hipStream_t st = h->stream;
    // doubles: keys in/out (2n), statistics (6); ints: flag, point flag, core,
    // the two lists (5n), counters (4)
    SHP_WS(buf, char, h, WS_FMM_OUTL, sizeof(double) * (2 * (size_t)n + 8) + sizeof(int) * (5 * (size_t)n + 8));
    double* kin = (double*)buf;
    double* kout = kin + n;
    double* stat = kout + n;
    int* flag = (int*)(stat + 8);
    int* pflag = flag + n;
    int* core = pflag + n;
    int* olist = core + n;
    int* plist = olist + n;
    int* cnt = plist + n;  // nvalid, nout, npoints
    hipLaunchKernelGGL(k_fmm_ifill, dim3(1), dim3(64), 0, st, cnt, 4, 0);
    size_t tsz = 0;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, tsz, kin, kout, n, 0, 64, st));
    const double* V[3] = {pt->x, pt->y, pt->z};
    for (int c = 0; c < 3; c++) {
        SHP_WS(cub, char, h, WS_CUB, tsz);
        hipLaunchKernelGGL(k_outl_keys, dim3(fdiv_up(n, 256)), dim3(256), 0, st, n, V[c], pt->status,
                           (const double*)nullptr, kin, c == 0 ? cnt : (int*)nullptr);
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(cub, tsz, kin, kout, n, 0, 64, st));
        hipLaunchKernelGGL(k_outl_median, dim3(1), dim3(64), 0, st, kout, cnt, 1.0, stat + c);
        hipLaunchKernelGGL(k_outl_keys, dim3(fdiv_up(n, 256)), dim3(256), 0, st, n, V[c], pt->status,
                           (const double*)(stat + c), kin, (int*)nullptr);
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(cub, tsz, kin, kout, n, 0, 64, st));
        hipLaunchKernelGGL(k_outl_median, dim3(1), dim3(64), 0, st, kout, cnt, kOutlierFactor, stat + 3 + c);
    }
    hipLaunchKernelGGL(k_outl_flags, dim3(fdiv_up(n, 256)), dim3(256), 0, st, n, pt->x, pt->y, pt->z, pt->status,
                       stat, flag, pflag, core);
    hipcub::CountingInputIterator<int> it(0);
    size_t t2 = 0;
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, t2, it, flag, olist, cnt + 1, n, st));
    SHP_WS(cub2, char, h, WS_CUB, t2);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub2, t2, it, flag, olist, cnt + 1, n, st));
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub2, t2, it, pflag, plist, cnt + 2, n, st));
    SHP_HIP_CHECK(hipGetLastError());
    int nout = 0;
    SHP_HIP_CHECK(hipMemcpyAsync(&nout, cnt + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    SHP_HIP_CHECK(hipStreamSynchronize(st));
    *o = OutlierSplit{flag, core, olist, plist, cnt + 2, nout};
    return SHP_OK;
}

int shp_fmm_dev(shp_handle h, const shp_particles* pt, int order, int leaf_level, int leaf_size,
                int ring, double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && pt && ax && ay && az, "shp_fmm_dev: null argument");
    const bool outliers = h->fmm_type == SHP_FMM_TREECODE_OUTLIERS || h->fmm_type == SHP_FMM_FMM_OUTLIERS;
    if (!outliers || pt->n <= 0) return fmm_eval(h, pt, order, leaf_level, leaf_size, ring, false, ax, ay, az);
    SHP_REQUIRE(pt->x && pt->y && pt->z && pt->q && pt->status && pt->species, "shp_fmm_dev: null particle array");
    if (order > kM2PMaxP) {
        set_error("shp_fmm_dev: outlier modes evaluate multipoles up to order %d (order %d)", kM2PMaxP, order);
        return SHP_ERR_UNSUPPORTED;
    }
    SHP_HIP_CHECK(hipSetDevice(h->device));
    OutlierSplit o;
    int rc = fmm_outliers(h, pt, &o);
    if (rc) return rc;
    // the tree holds the other ions; the outliers get its field through M2P
    // and their near cells, and every target gets the outliers' field directly
    shp_particles core = *pt;
    core.status = o.core;
    const FmmPoints P{o.plist, o.pcount, o.nout};
    rc = fmm_eval(h, &core, order, leaf_level, leaf_size, ring, false, ax, ay, az, &P);
    if (rc) return rc;
    if (o.nout > 0)
        hipLaunchKernelGGL(k_outl_sources, dim3(fdiv_up(pt->n, 256)), dim3(256), 0, h->stream, pt->n, pt->x, pt->y,
                           pt->z, pt->q, pt->status, pt->species, o.olist, o.nout, k_const_host(), ax, ay, az);
    SHP_HIP_CHECK(hipGetLastError());
    return SHP_OK;
}

int shp_fmm_set_type(shp_handle h, int fmm_type) {
    SHP_REQUIRE(h, "shp_fmm_set_type: null handle");
    SHP_REQUIRE(fmm_type >= SHP_FMM_TREECODE && fmm_type <= SHP_FMM_FMM_OUTLIERS, "shp_fmm_set_type: type %d",
                fmm_type);
    h->fmm_type = fmm_type;
    return SHP_OK;
}

int shp_fmm_outliers_dev(shp_handle h, const shp_particles* pt, int* flags, int* nout) {
    SHP_REQUIRE(h && pt && flags, "shp_fmm_outliers_dev: null argument");
    if (nout) *nout = 0;
    if (pt->n <= 0) return SHP_OK;
    SHP_REQUIRE(pt->x && pt->y && pt->z && pt->status, "shp_fmm_outliers_dev: null particle array");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    OutlierSplit o;
    const int rc = fmm_outliers(h, pt, &o);
    if (rc) return rc;
    SHP_HIP_CHECK(hipMemcpyAsync(flags, o.flag, sizeof(int) * pt->n, hipMemcpyDeviceToDevice, h->stream));
    if (nout) *nout = o.nout;
    return SHP_OK;
}

int shp_fmm_shell_orders(int order, int ring, int sym_type, int* out3) {
    SHP_REQUIRE(out3, "shp_fmm_shell_orders: null argument");
    m2l_orders(order, ring, sym_type == SHP_SYM_ROTATIONAL, out3[0], out3[1], out3[2]);
    return SHP_OK;
}

int shp_fmm_last_leaf_level(shp_handle h) { return h ? last_leaf_level(h) : -1; }

int shp_fmm_level_estimates(shp_handle h, double* out, int cap) {
    SHP_REQUIRE(h && (out || cap == 0), "shp_fmm_level_estimates: null argument");
    const int n = std::min(cap, h->fmm_est_n);
    for (int k = 0; k < 4 * n; k++) out[k] = h->fmm_est[k];
    return h->fmm_est_n;
}

int shp_fmm_stats(shp_handle h, double* out) {
    SHP_REQUIRE(h && out, "shp_fmm_stats: null argument");
    SHP_REQUIRE(h->fmm_have_last || h->fmm_pending, "shp_fmm_stats: no FMM evaluation on this handle yet");
    const int rc = fmm_fold(h);
    if (rc != SHP_OK) return rc;
    for (int k = 0; k < kFmmStats; k++) out[k] = h->fmm_last[k];
    return SHP_OK;
}

int shp_fmm_timing(shp_handle h, double* out, int reset) {
    SHP_REQUIRE(h, "shp_fmm_timing: null handle");
    const int rc = fmm_fold(h);
    if (rc != SHP_OK) return rc;
    if (out) {
        for (int k = 0; k < kFmmStats; k++) out[k] = h->fmm_acc[k];
        out[kFmmStats] = (double)h->fmm_acc_n;
    }
    if (reset) {
        for (int k = 0; k < kFmmStats; k++) h->fmm_acc[k] = 0.0;
        h->fmm_acc_n = 0;
    }
    return SHP_OK;
}

int shp_fmm_sym_dev(shp_handle h, const shp_particles* pt, const shp_symmetry* sym, int order,
                    int leaf_level, int leaf_size, int ring, double* ax, double* ay, double* az) {
    SHP_REQUIRE(h && pt && ax && ay && az, "shp_fmm_sym_dev: null argument");
    const int symt = sym ? sym->type : SHP_SYM_NONE;
    if (symt == SHP_SYM_NONE) return shp_fmm_dev(h, pt, order, leaf_level, leaf_size, ring, ax, ay, az);
    if (h->fmm_type == SHP_FMM_TREECODE_OUTLIERS || h->fmm_type == SHP_FMM_FMM_OUTLIERS) {
        set_error("shp_fmm_sym_dev: the outlier modes have no symmetry images here");
        return SHP_ERR_UNSUPPORTED;
    }
    const int n = pt->n;
    if (n <= 0) return SHP_OK;
    SHP_REQUIRE(pt->x && pt->y && pt->z && pt->q && pt->status && pt->species,
                "shp_fmm_sym_dev: null particle array");
    std::vector<double4> im;
    im.push_back(make_double4(1.0, 0.0, 0.0, 0.0));
    if (symt == SHP_SYM_ROTATIONAL) {
        SHP_REQUIRE(sym->nsec >= 1 && sym->nsec <= 64, "shp_fmm_sym_dev: nsec=%d", sym->nsec);
        for (int k = 1; k < sym->nsec; k++) {
            double c, sn;
            if ((4 * k) % sym->nsec == 0) {  // exact quadrant rotations
                const int qd = (4 * k / sym->nsec) & 3;
                const int cs[4] = {1, 0, -1, 0};
                c = cs[qd];
                sn = cs[(qd + 3) & 3];
            } else {
                const double ang = 2.0 * 3.14159265358979323846 * k / sym->nsec;
                c = cos(ang);
                sn = sin(ang);
            }
            im.push_back(make_double4(c, sn, 0.0, 0.0));
        }
    } else {
        SHP_REQUIRE(symt == SHP_SYM_TRANSLATIONAL && sym->nimg >= 1 && sym->lattice,
                    "shp_fmm_sym_dev: translational symmetry needs a lattice");
        bool centre = false;
        for (int k = 0; k < sym->nimg; k++) {
            const int* l = sym->lattice + 3 * k;
            if (l[0] == 0 && l[1] == 0 && l[2] == 0) {
                centre = true;
                continue;
            }
            im.push_back(make_double4(l[0] * sym->domain[0], l[1] * sym->domain[1],
                                      l[2] * sym->domain[2], 1.0));
        }
        SHP_REQUIRE(centre, "shp_fmm_sym_dev: the lattice must contain the central cell (0,0,0)");
    }
    const int nblk = (int)im.size();
    SHP_REQUIRE((size_t)n * nblk < ((size_t)1 << 31), "shp_fmm_sym_dev: %d ions x %d images too many", n, nblk);
    const int ntot = n * nblk;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(dimg, double4, h, WS_FMM_SYM_IMG, nblk);
    SHP_WS(dd, double, h, WS_FMM_SYM_D, 7 * (size_t)ntot);
    SHP_WS(di, int, h, WS_FMM_SYM_I, 2 * (size_t)ntot);
    SHP_HIP_CHECK(hipMemcpyAsync(dimg, im.data(), sizeof(double4) * nblk, hipMemcpyHostToDevice, h->stream));
    double *X2 = dd, *Y2 = dd + ntot, *Z2 = dd + 2 * (size_t)ntot, *Q2 = dd + 3 * (size_t)ntot;
    double *A2 = dd + 4 * (size_t)ntot;
    int *st2 = di, *sp2 = di + ntot;
    hipLaunchKernelGGL(k_fmm_images, dim3(fdiv_up(ntot, 256)), dim3(256), 0, h->stream, n, nblk, dimg, pt->x,
                       pt->y, pt->z, pt->q, pt->status, pt->species, X2, Y2, Z2, Q2, st2, sp2);
    SHP_HIP_CHECK(hipGetLastError());
    shp_particles all{ntot, X2, Y2, Z2, Q2, st2, sp2};
    int rc = fmm_eval(h, &all, order, leaf_level, leaf_size, ring, symt == SHP_SYM_ROTATIONAL, A2, A2 + ntot,
                      A2 + 2 * (size_t)ntot);
    if (rc) return rc;
    SHP_HIP_CHECK(hipMemcpyAsync(ax, A2, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(ay, A2 + ntot, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(az, A2 + 2 * (size_t)ntot, sizeof(double) * n, hipMemcpyDeviceToDevice,
                                 h->stream));
    return SHP_OK;
}

}  // extern "C"
