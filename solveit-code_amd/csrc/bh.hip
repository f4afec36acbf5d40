// Barnes-Hut force on gfx950: the reference's grid of root cells and octree,
// built on the GPU from Morton-sorted ions, stackless threaded walk, reference
// (FP32) arithmetic.
//
// Replaces BarnesHutForceCalculation::updateAcceleration and its passes
// (barnes_hut_force_calculation.cpp:574-715: computeWorldLimits :456-522,
// generateTree :258-331, locateSubcell :149-216, addToCell :224-253,
// computeCellProperties :337-390, computeForceFromTree :417-450,
// barnesHutKernelCalculation :539-569), DomainBoundsFunction::computeDomainBounds
// (domain_bounds_functions.cpp:121-316), the rotational
// (barnes_hut_force_calculation_rotational.cpp:28-135) and translational
// (barnes_hut_force_calculation_translational.cpp:190-280) variants, and
// cell.cpp:7-131 for the cell geometry.
//
// Root cells.  The float bounding box of the sources, seeded with the first
// ion of the range whatever its status, is reduced on the GPU; the host then
// lays the reference's grid of root cells from it (cubic cells by default:
// the smallest non-zero extent is the first cell size, the other axes are cut
// into floor/ceil(extent / size) cells, the size doubles until a cell holds
// 100 ions on average, is scaled by 1 + 0.01 / n and centred; cells are then
// added until every source is covered and merged 2x2x2 while the smallest
// count is >= 2).  Rectangular-prism root cells and forced cell counts
// (setDomainToggle, setBoxDimVals) are options.  An ion goes to the root cell
// (pos - roots) / dims truncates to, or to the last cell of the chain when
// none matches, as generateTree does.
//
// Tree.  The reference inserts ions one by one into a pointer tree and splits
// a cell as soon as it holds two ions, so its cells are exactly the cells
// holding two or more ions plus their children, independent of the insertion
// order.  Here every source ion gets its root cell and 42 octant digits (two
// 64-bit words) by descending from that root cell with the reference's own
// float compares against the cell centres (location + dims / 2, children at
// location + bit * dims / 2); the (root, digits) keys are radix-sorted and the
// tree is built level by level: a node is a run of sorted keys with a common
// prefix; a node of two or more ions is split into its non-empty children (8
// binary searches), which are allocated contiguously in octant order with the
// threaded successor `next` (next sibling, or the parent's successor).  The
// non-empty root cells form level 0, chained in the reference's x-y-z order.
// Empty cells of the reference carry no charge and are skipped by its walk,
// so omitting them changes no sum.  Ions that share all 42 digits (closer
// than 2^-42 of a root cell: coincident ions, for which the reference's
// insertion does not terminate) form one bucket leaf.
//
// Cell properties are formed bottom-up in the reference's order (children in
// octant order): net charge, sum |q|, centre of charge = sum centre |net q| /
// sum |q| (geometric centre when the sum is zero), delta = |centre of charge
// - geometric centre|.  The walk is the reference's: leaves interact
// directly, an inner cell is accepted iff theta (d - delta) > repLength with d
// the softened float distance, otherwise opened; the successor pointer makes
// the walk stackless and runs over the whole chain of root cells.  All float
// arithmetic follows the reference operation by operation (no contraction),
// so the result equals the reference algorithm's.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "shp_internal.h"

#pragma clang fp contract(off)

namespace shp {
namespace {

constexpr int kBhDepth = 42;           // levels below a root cell carried by the keys
constexpr int kBhWord = 21;            // octant digits per 64-bit key word
constexpr int kBhThreads = 256;
constexpr int kBhMinAvgPPB = 100;      // FMMConstant::minAvgPPB (constants.h:80)
constexpr float kBhDomainEps = 0.01f;  // DomainBoundsFunction::domainEps (domain_bounds_functions.h:39)
constexpr int kBhMaxRoots = 1 << 24;
constexpr long kBhCoherentMin = 1 << 18;  // sources from which the wave-coherent walk is used (walk = 0)

inline int bdiv(long a, long b) { return (int)((a + b - 1) / b); }

// The grid of root cells: location of the first cell, cell dimensions, cells
// per axis (x outermost, z innermost in the chain).
struct BhLayout {
    float roots[3];
    float dims[3];
    int nd[3];
};

// d[l] = (dx, dy, dz, repLength) of the cells l levels below a root cell;
// each level halves the one above (Cell::generateSubcells, cell.cpp:114-131).
struct BhLevels {
    float4 d[kBhDepth + 2];
};

// Small header copied to the host once per call.
struct BhHead {
    int ns;
    int pad;
    float mn[3];
    float mx[3];
};

struct BhNodes {
    int* start;
    int* count;
    int* first;   // first child (-1: leaf)
    int* nch;
    int* next;    // threaded successor (-1: end of walk)
    float4* geo;  // location (corner) xyz, repLength
    float4* cq;   // centre of charge xyz, net charge
    float2* du;   // delta, sum |q|
};

// What the walk reads of a node, in one 32-byte record, with the nodes in
// depth-first preorder (the order of the reference's walk): the centre of
// charge and net charge, then (1 for a cell / -1 for a leaf, the size of its
// subtree, and for a cell delta and repLength as float bits, for a leaf the
// ion count and first ion).  In preorder a cell's first child is the next
// record and its successor lies `size` records on.
struct BhWalk {
    float4* cq;
    int4* aux;
};

// float -> int as x86-64's cvttss2si (the reference's platform): truncation,
// INT_MIN for NaN and out-of-range values.
__host__ __device__ inline int bh_f2i(float v) {
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT_MIN;
}

__host__ __device__ inline float bh_max3(float a, float b, float c) {
    float m = a;
    if (b > m) m = b;
    if (c > m) m = c;
    return m;
}

}  // namespace

// computeDomainBounds (outlier flags all 0) from the sources' float bounds
// (mn, mx), the first ion of the range and the number of sources:
// boxMin, the cell dimensions and the cells per axis, plus the bounds seeded
// with the first ion (minO, maxO).  toggle 0: cubic cells, 1: rectangular prism.
// rot != 0: computeDomainBoundsRotations (domain_bounds_functions.cpp:317-500),
// whose x, y extent is the square [-R, R]^2, R the larger of the x and y maxima.
void domain_bounds_host(const float mn_src[3], const float mx_src[3], const float seed[3], int nsrc,
                        int toggle, int rot, const int* forced, float boxMin[3], float dims[3], int bdv[3],
                        float minO[3], float maxO[3]) {
    float rootmax[3];
    for (int c = 0; c < 3; c++) {
        boxMin[c] = minO[c] = mn_src[c] < seed[c] ? mn_src[c] : seed[c];
        rootmax[c] = maxO[c] = mx_src[c] > seed[c] ? mx_src[c] : seed[c];
    }
    const float subtr = 0.0f;
    if (rot) {
        rootmax[0] = rootmax[0] < rootmax[1] ? rootmax[1] : rootmax[0];
        rootmax[1] = rootmax[0];
        boxMin[0] = -rootmax[0];
        boxMin[1] = -rootmax[1];
        dims[0] = rootmax[0] * 2;
        dims[1] = rootmax[1] * 2;
        dims[2] = rootmax[2] - boxMin[2];
    } else {
        for (int c = 0; c < 3; c++) {
            boxMin[c] = boxMin[c] - subtr;
            dims[c] = (rootmax[c] - boxMin[c]) + subtr;
        }
    }
    bdv[0] = bdv[1] = bdv[2] = 1;
    if (toggle == 0) {
        float dd[3] = {dims[0], dims[1], dims[2]};
        int di[3] = {0, 1, 2};
        for (int i = 0; i < 3; i++)  // ascending; ties keep the axis order
            for (int j = i; j > 0 && dd[j - 1] > dd[j]; j--) {
                std::swap(dd[j - 1], dd[j]);
                std::swap(di[j - 1], di[j]);
            }
        float size = 0.0f;
        int i = -1;
        do {
            i++;
            size = dd[i];
            bdv[di[i]] = 1;
        } while (size == 0.0f && i < 2);
        const float soft = (float)kSoftening;
        size = size < soft ? soft : size;
        for (;;) {
            for (int j = i + 1; j < 3; j++) {
                const float ratio = dd[j] / size;
                if ((double)ratio - std::floor((double)ratio) >= 0.5) {
                    bdv[di[j]] = (int)std::ceil((double)ratio);
                } else {
                    if (ratio < 2.0f) bdv[di[j]] = j == 0 ? 1 : bdv[di[j - 1]];
                    else bdv[di[j]] = (int)std::floor((double)ratio);
                    if ((double)ratio > 0.5 || j == 2) size = dd[j] / (float)bdv[di[j]];
                }
            }
            const int ndom = bdv[0] * bdv[1] * bdv[2];
            const int avg = (nsrc + ndom - 1) / ndom;
            if (size == dd[2] || avg >= kBhMinAvgPPB || size == soft) break;
            size *= 2.0f;
        }
        if (forced) {
            for (int c = 0; c < 3; c++) bdv[c] = forced[c];
            size = dims[0] / (float)bdv[0];
            float r = dims[1] / (float)bdv[1];
            size = size < r ? r : size;
            r = dims[2] / (float)bdv[2];
            size = size < r ? r : size;
            size = size < soft ? soft : size;
            if (bdv[0] >= bdv[1] && bdv[0] >= bdv[2]) di[2] = 0;
            else if (bdv[1] >= bdv[0] && bdv[1] >= bdv[2]) di[2] = 1;
            else if (bdv[2] >= bdv[0] && bdv[2] >= bdv[1]) di[2] = 2;
        }
        size *= 1.0f + kBhDomainEps / (float)bdv[di[2]];
        for (int c = 0; c < 3; c++) {
            boxMin[c] = ((rootmax[c] + subtr) + boxMin[c] - size * (float)bdv[c]) / 2.0f;
            dims[c] = size;
        }
    }
}

namespace {

// computeDomainBounds followed by computeWorldLimits.
void bh_layout_host(const float mn_src[3], const float mx_src[3], const float seed[3], int nsrc,
                    int toggle, const int* forced, BhLayout* L) {
    float boxMin[3], minO[3], maxO[3], dims[3];
    int bdv[3];
    domain_bounds_host(mn_src, mx_src, seed, nsrc, toggle, 0, forced, boxMin, dims, bdv, minO, maxO);
    float roots[3] = {boxMin[0], boxMin[1], boxMin[2]};
    for (int c = 0; c < 3; c++)
        while (roots[c] > minO[c]) {
            roots[c] -= dims[c];
            bdv[c]++;
        }
    for (int c = 0; c < 3; c++)
        while (roots[c] + (float)bdv[c] * dims[c] < maxO[c]) bdv[c]++;
    if (toggle == 0) {
        float size = dims[0];
        int dc = std::min(bdv[0], std::min(bdv[1], bdv[2]));
        while (dc / 2 > 0) {
            dc = (dc + 1) / 2;
            size *= 2.0f;
            for (int c = 0; c < 3; c++) bdv[c] = (bdv[c] + 1) / 2;
        }
        for (int c = 0; c < 3; c++) dims[c] = size;
    }
    for (int c = 0; c < 3; c++) {
        L->roots[c] = roots[c];
        L->dims[c] = dims[c];
        L->nd[c] = bdv[c];
    }
}

void bh_levels_host(const BhLayout& L, BhLevels* lv) {
    float d[3] = {L.dims[0], L.dims[1], L.dims[2]};
    for (int l = 0; l < kBhDepth + 2; l++) {
        lv->d[l] = make_float4(d[0], d[1], d[2], bh_max3(d[0], d[1], d[2]));
        for (int c = 0; c < 3; c++) d[c] = d[c] / 2.0f;
    }
}

// Location of root cell r of the chain (generateTree :263-277).
__host__ __device__ inline float3 bh_root_loc(const BhLayout& L, int r) {
    const int k = r % L.nd[2];
    const int j = (r / L.nd[2]) % L.nd[1];
    const int i = r / (L.nd[2] * L.nd[1]);
    if (r == 0) return make_float3(L.roots[0], L.roots[1], L.roots[2]);
    return make_float3(L.roots[0] + (float)i * L.dims[0], L.roots[1] + (float)j * L.dims[1],
                       L.roots[2] + (float)k * L.dims[2]);
}

// Subtree sizes, one level at a time from the deepest (children are
// contiguous: first .. first + nch - 1).
__global__ void k_bh_size(BhNodes t, int lev, int* __restrict__ size, const int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    int sz = 1;
    const int f = t.first[i];
    if (f >= 0)
        for (int c = f; c < f + t.nch[i]; c++) sz += size[c];
    size[i] = sz;
}

// Preorder positions of the children of the level-lev cells, from the top:
// the first child follows its parent, each next sibling follows the subtree
// of the one before (the root cells get theirs from a scan of their sizes).
__global__ void k_bh_pre(BhNodes t, int lev, const int* __restrict__ size, int* __restrict__ pre,
                         const int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    const int f = t.first[i];
    if (f < 0) return;
    int p = pre[i] + 1;
    for (int c = f; c < f + t.nch[i]; c++) {
        pre[c] = p;
        p += size[c];
    }
}

__global__ void k_bh_pack(BhNodes t, BhWalk w, const int* __restrict__ size, const int* __restrict__ pre,
                          const int* __restrict__ misc) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= misc[0]) return;
    const int k = pre[node];
    w.cq[k] = t.cq[node];
    w.aux[k] = t.first[node] < 0
                   ? make_int4(-1, 1, t.count[node], t.start[node])
                   : make_int4(1, size[node], __float_as_int(t.du[node].x), __float_as_int(t.geo[node].w));
}

// ---------------------------------------------------------------- inputs ----

__global__ void k_bh_from_f64(const double* __restrict__ X, const double* __restrict__ Y,
                              const double* __restrict__ Z, const double* __restrict__ Q, int n,
                              float4* __restrict__ P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) P[i] = make_float4((float)X[i], (float)Y[i], (float)Z[i], (float)Q[i]);
}

__global__ void k_bh_from_aos(const float4* __restrict__ vel4, const float4* __restrict__ acc4, int n,
                              int* __restrict__ st, int* __restrict__ sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float s = acc4[i].w;
        st[i] = s == 1.0f ? 1 : (s == 2.0f ? 2 : (s == 0.0f ? 0 : -1));
        sp[i] = (int)vel4[i].w;
    }
}

__global__ void k_bh_flags(const int* __restrict__ st, int i0, int m, char* __restrict__ flags) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < m) {
        const int s = st[i0 + t];
        flags[t] = (s == 1 || s == 2) ? 1 : 0;
    }
}

// ---------------------------------------------------------------- bounds ----

// Float min / max of the sources, per block.  min and max are exact, so the
// order of the reduction does not matter.
__global__ __launch_bounds__(kBhThreads) void k_bh_bbox_partial(const float4* __restrict__ P,
                                                                const int* __restrict__ src,
                                                                const int* __restrict__ nsrc,
                                                                float* __restrict__ part) {
    __shared__ float red[6][kBhThreads];
    const int ns = *nsrc;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) {
        const float4 p = P[src[k]];
        const float v[3] = {p.x, p.y, p.z};
        for (int c = 0; c < 3; c++) {
            mn[c] = v[c] < mn[c] ? v[c] : mn[c];
            mx[c] = v[c] > mx[c] ? v[c] : mx[c];
        }
    }
    for (int c = 0; c < 3; c++) {
        red[c][threadIdx.x] = mn[c];
        red[3 + c][threadIdx.x] = mx[c];
    }
    __syncthreads();
    for (int s = kBhThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                const float a = red[c][threadIdx.x + s], b = red[3 + c][threadIdx.x + s];
                red[c][threadIdx.x] = a < red[c][threadIdx.x] ? a : red[c][threadIdx.x];
                red[3 + c][threadIdx.x] = b > red[3 + c][threadIdx.x] ? b : red[3 + c][threadIdx.x];
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Folds the block partials into the header (one wave).
__global__ __launch_bounds__(64) void k_bh_bbox_final(const float* __restrict__ part, int nparts,
                                                      const int* __restrict__ nsrc,
                                                      BhHead* __restrict__ head) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nparts; b += 64)
        for (int c = 0; c < 3; c++) {
            mn[c] = part[b * 6 + c] < mn[c] ? part[b * 6 + c] : mn[c];
            mx[c] = part[b * 6 + 3 + c] > mx[c] ? part[b * 6 + 3 + c] : mx[c];
        }
    for (int off = 32; off > 0; off >>= 1)
        for (int c = 0; c < 3; c++) {
            const float a = __shfl_xor(mn[c], off), b = __shfl_xor(mx[c], off);
            mn[c] = a < mn[c] ? a : mn[c];
            mx[c] = b > mx[c] ? b : mx[c];
        }
    if (threadIdx.x == 0) {
        head->ns = *nsrc;
        for (int c = 0; c < 3; c++) {
            head->mn[c] = mn[c];
            head->mx[c] = mx[c];
        }
    }
}

// ------------------------------------------------------------------ keys ----

// Root cell (rk) and 42 octant digits (kb: levels 1-21, kc: levels 22-42;
// x -> bit 0, y -> bit 1, z -> bit 2 of a digit; the first level of a word
// in its top digit) per source, from the reference's float compares.
__global__ void k_bh_keys(const float4* __restrict__ P, const int* __restrict__ src, int ns, BhLayout L,
                          BhLevels lv, unsigned* __restrict__ rk, uint64_t* __restrict__ kb,
                          uint64_t* __restrict__ kc, int* __restrict__ vals,
                          unsigned long long* __restrict__ unmatched) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const float4 p = P[src[k]];
    const int nroot = L.nd[0] * L.nd[1] * L.nd[2];
    const int di = bh_f2i((p.x - L.roots[0]) / L.dims[0]);
    const int dj = bh_f2i((p.y - L.roots[1]) / L.dims[1]);
    const int dk = bh_f2i((p.z - L.roots[2]) / L.dims[2]);
    int r;
    if (di < 0 || di >= L.nd[0] || dj < 0 || dj >= L.nd[1] || dk < 0 || dk >= L.nd[2]) {
        r = nroot - 1;
        atomicAdd(unmatched, 1ull);
    } else {
        r = (di * L.nd[1] + dj) * L.nd[2] + dk;
    }
    const float3 loc = bh_root_loc(L, r);
    float lx = loc.x, ly = loc.y, lz = loc.z;
    uint64_t w0 = 0, w1 = 0;
#pragma unroll
    for (int lev = 0; lev < kBhDepth; lev++) {
        const float4 h = lv.d[lev + 1];
        const int bx = p.x > lx + h.x, by = p.y > ly + h.y, bz = p.z > lz + h.z;
        lx += (float)bx * h.x;
        ly += (float)by * h.y;
        lz += (float)bz * h.z;
        const uint64_t dg = (uint64_t)(bx | (by << 1) | (bz << 2));
        if (lev < kBhWord) w0 = (w0 << 3) | dg;
        else w1 = (w1 << 3) | dg;
    }
    rk[k] = (unsigned)r;
    kb[k] = w0;
    kc[k] = w1;
    vals[k] = k;
}

template <typename T>
__global__ void k_bh_permute(const T* __restrict__ in, const int* __restrict__ perm, int n,
                             T* __restrict__ out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) out[s] = in[perm[s]];
}

__global__ void k_bh_gather(const float4* __restrict__ P, const int* __restrict__ src,
                            const int* __restrict__ order, int ns, float4* __restrict__ SP,
                            int* __restrict__ SI) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int i = src[order[s]];
    SP[s] = P[i];
    SI[s] = i;
}

// ------------------------------------------------------------------ tree ----

// misc[0] = nodes allocated, misc[1] = overflow flag, misc[2 + l] = first node
// of level l (l = 0 .. kBhDepth + 1).  Level 0: the non-empty root cells in
// chain order (uniq, cnt from the run-length encoding of the sorted root
// indices rkS).
__global__ void k_bh_init(BhNodes t, const unsigned* __restrict__ uniq, const int* __restrict__ cnt,
                          const int* __restrict__ nruns, const unsigned* __restrict__ rkS, int ns,
                          BhLayout L, float rep0, int* __restrict__ misc) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int nr = *nruns;
    if (r == 0) {
        misc[0] = nr;
        misc[1] = 0;
        misc[2] = 0;
        misc[3] = nr;
    }
    if (r >= nr) return;
    const unsigned root = uniq[r];
    int lo = 0, hi = ns;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rkS[mid] < root) lo = mid + 1;
        else hi = mid;
    }
    const float3 loc = bh_root_loc(L, (int)root);
    t.start[r] = lo;
    t.count[r] = cnt[r];
    t.first[r] = -1;
    t.nch[r] = 0;
    t.next[r] = r + 1 < nr ? r + 1 : -1;
    t.geo[r] = make_float4(loc.x, loc.y, loc.z, rep0);
}

// Splits the nodes of level lev (< kBhDepth) holding two or more ions into
// their non-empty children, from the digit of level lev + 1.
__global__ void k_bh_split(BhNodes t, int lev, const uint64_t* __restrict__ key, int shift, float4 h,
                           int cap, int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    const int c = t.count[i];
    if (c < 2) return;
    const int s = t.start[i], e = s + c;
    int b[9];
    b[0] = s;
    b[8] = e;
    for (int d = 1; d < 8; d++) {  // first key whose digit is >= d
        int lo = b[d - 1], hi = e;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)((key[mid] >> shift) & 7) < d) lo = mid + 1;
            else hi = mid;
        }
        b[d] = lo;
    }
    int ne = 0;
    for (int d = 0; d < 8; d++) ne += b[d + 1] > b[d];
    const int base = atomicAdd(&misc[0], ne);
    if (base + ne > cap) {
        atomicOr(&misc[1], 1);
        return;
    }
    const float4 g = t.geo[i];
    const int nx = t.next[i];
    int j = 0;
    for (int d = 0; d < 8; d++) {
        if (b[d + 1] == b[d]) continue;
        const int k = base + j;
        t.start[k] = b[d];
        t.count[k] = b[d + 1] - b[d];
        t.first[k] = -1;
        t.nch[k] = 0;
        t.next[k] = j + 1 < ne ? k + 1 : nx;
        t.geo[k] = make_float4(g.x + (float)(d & 1) * h.x, g.y + (float)((d >> 1) & 1) * h.y,
                               g.z + (float)((d >> 2) & 1) * h.z, h.w);
        j++;
    }
    t.first[i] = base;
    t.nch[i] = ne;
}

// Level lev + 1 ends where the allocation stands after splitting level lev.
__global__ void k_bh_mark(int lev, int cap, int* __restrict__ misc) {
    if (threadIdx.x == 0) misc[2 + lev + 2] = min(misc[0], cap);
}

// h = half the dimensions of a level-lev cell (the dimensions of level lev + 1).
__global__ void k_bh_props(BhNodes t, int lev, float4 h, const float4* __restrict__ SP,
                           const int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    const float4 g = t.geo[i];
    const float gx = g.x + h.x, gy = g.y + h.y, gz = g.z + h.z;
    float q = 0.0f, uq = 0.0f, cx, cy, cz;
    const int f = t.first[i];
    if (f < 0 && t.count[i] == 1) {
        const float4 p = SP[t.start[i]];
        q = p.w;
        uq = fabsf(p.w);
        cx = p.x;
        cy = p.y;
        cz = p.z;
    } else {
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        if (f < 0) {  // bucket at the deepest level: ions as children
            for (int j = t.start[i], e = j + t.count[i]; j < e; j++) {
                const float4 p = SP[j];
                q += p.w;
                uq += fabsf(p.w);
                tx += p.x * fabsf(p.w);
                ty += p.y * fabsf(p.w);
                tz += p.z * fabsf(p.w);
            }
        } else {
            for (int j = f, e = f + t.nch[i]; j < e; j++) {
                const float4 c = t.cq[j];
                q += c.w;
                uq += t.du[j].y;
                tx += c.x * fabsf(c.w);
                ty += c.y * fabsf(c.w);
                tz += c.z * fabsf(c.w);
            }
        }
        if (uq != 0.0f) {
            cx = tx / uq;
            cy = ty / uq;
            cz = tz / uq;
        } else {
            cx = gx;
            cy = gy;
            cz = gz;
        }
    }
    const float dx = cx - gx, dy = cy - gy, dz = cz - gz;
    t.cq[i] = make_float4(cx, cy, cz, q);
    t.du[i] = make_float2(sqrtf(dx * dx + dy * dy + dz * dz), uq);
}

// Tree summary for shp_bh_stats: deepest non-empty level and bucket leaves.
__global__ void k_bh_tree_stats(BhNodes t, const int* __restrict__ misc,
                                unsigned long long* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        int deepest = 0;
        for (int l = 0; l <= kBhDepth; l++)
            if (misc[2 + l + 1] > misc[2 + l]) deepest = l;
        out[0] = (unsigned long long)deepest;
    }
    const int lo = misc[2 + kBhDepth], hi = misc[2 + kBhDepth + 1];
    if (i < hi - lo && t.count[lo + i] >= 2) atomicAdd(&out[1], 1ull);
}

// ------------------------------------------------------------------ walk ----

// computeEffectiveDistance: float squares, softening and sqrt in double.
__device__ __forceinline__ float bh_dist(float px, float py, float pz, float cx, float cy, float cz,
                                         float& rx, float& ry, float& rz) {
    rx = px - cx;
    ry = py - cy;
    rz = pz - cz;
    const float s = rx * rx + ry * ry + rz * rz;
    return (float)sqrt((double)s + kSoftening);
}

template <int SYM>
__device__ __forceinline__ void bh_add(float f, float rx, float ry, float rz, int cb, int sb, float& ax,
                                       float& ay, float& az) {
    if (SYM == SHP_SYM_ROTATIONAL) {
        ax += f * ((float)cb * rx - (float)sb * ry);
        ay += f * ((float)sb * rx + (float)cb * ry);
    } else {
        ax += f * rx;
        ay += f * ry;
    }
    az += f * rz;
}

// quadrantRotationFunction, auxiliary_functions.h:55-57
__device__ __forceinline__ int bh_qrf(int n) {
    const int oe = (n & 1) ? -1 : 1;
    const int a = 1 - oe;
    const int sg = (a > 0) - (a < 0);
    const int oe2 = (((n - 1) >> 1) & 1) ? -1 : 1;
    return sg * oe2;
}

// The walk of one target image: its (transformed) position, its sums, its
// next record, and whether its own ion is skipped.
struct BhPath {
    float px, py, pz, ax, ay, az;
    int node, s;
    bool self;
};

// One step of a walk at record k = (q, x): a bucket of coincident ions (rare),
// else one distance for a single-ion leaf (its centre of charge is the ion)
// or a cell, the reference's operations per ion in either case.  Returns the
// next record: the following one (leaf, opened cell) or past the subtree
// (accepted cell).
template <int SYM>
__device__ __forceinline__ int bh_step(const float4* __restrict__ SP, float theta, int cb, int sb, int k,
                                       const float4 q, const int4 x, BhPath& w, unsigned& ncell, unsigned& nion,
                                       unsigned& nopen) {
    const bool leaf = x.x < 0;
    if (leaf && x.z != 1) {
        for (int j = x.w; j < x.w + x.z; j++) {
            if (w.self && j == w.s) continue;
            const float4 qj = SP[j];
            float rx, ry, rz;
            const float d = bh_dist(w.px, w.py, w.pz, qj.x, qj.y, qj.z, rx, ry, rz);
            bh_add<SYM>(qj.w / (d * d * d), rx, ry, rz, cb, sb, w.ax, w.ay, w.az);
            nion++;
        }
        return k + 1;
    }
    float rx, ry, rz;
    const float d = bh_dist(w.px, w.py, w.pz, q.x, q.y, q.z, rx, ry, rz);
    const bool take = leaf ? !(w.self && x.w == w.s) : theta * (d - __int_as_float(x.z)) > __int_as_float(x.w);
    if (take) {
        bh_add<SYM>(q.w / (d * d * d), rx, ry, rz, cb, sb, w.ax, w.ay, w.az);
        if (leaf) nion++;
        else ncell++;
    }
    if (!leaf && !take) nopen++;
    return (!leaf && take) ? k + x.y : k + 1;
}

__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off));
    return v;
}

// One thread per source in key order; the targets (status 1) walk the tree.
// The lanes of a wave walk together in preorder: each step visits the
// smallest next record over the lanes, which only the lanes waiting for it
// process; a lane that accepted a cell waits past its subtree.  Every lane
// so visits its own nodes in its own order, as the reference's walk does, and
// each record is one wave-uniform (scalar) load instead of 64 scattered ones.
// ctr[0] += accepted cells, ctr[1] += ion terms, ctr[3] += opened cells (one
// wave-reduced atomic each per wave).  With COHERENT false every lane walks
// on its own (the same steps, in the same order, with its own record loads):
// faster when there are too few waves to hide the coherent walk's chain of
// wave-wide minimum, scalar load and step.
template <int SYM, bool COHERENT>
__global__ __launch_bounds__(kBhThreads) void k_bh_walk(
    BhWalk t, int nrec, const float4* __restrict__ SP, const int* __restrict__ SI, int ns,
    const int* __restrict__ st, const int* __restrict__ sp, float theta, const int* __restrict__ lat,
    int nimg, float dx0, float dy0, float dz0, double K, int i0, double* __restrict__ ax_out,
    double* __restrict__ ay_out, double* __restrict__ az_out, unsigned long long* __restrict__ ctr) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = s < ns ? SI[s] : 0;
    const bool active = s < ns && st[i] == 1;
    unsigned ncell = 0, nion = 0, nopen = 0;
    const float4 p = active ? SP[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const double kc = active ? K * (double)p.w / species_mass(sp[i]) : 0.0;
    const int nrep = SYM == SHP_SYM_ROTATIONAL ? 4 : (SYM == SHP_SYM_TRANSLATIONAL ? nimg : 1);
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (int k = 0; k < nrep; k++) {
        BhPath w{p.x, p.y, p.z, 0.0f, 0.0f, 0.0f, active ? 0 : nrec, s, true};
        int cb = 1, sb = 0;
        if (SYM == SHP_SYM_ROTATIONAL) {
            const int cf = bh_qrf(5 - k), sf = bh_qrf(4 - k);
            cb = bh_qrf(k + 1);
            sb = bh_qrf(k);
            w.px = (float)cf * p.x - (float)sf * p.y;
            w.py = (float)sf * p.x + (float)cf * p.y;
            w.self = k == 0;
        } else if (SYM == SHP_SYM_TRANSLATIONAL) {
            const int lx = lat[3 * k], ly = lat[3 * k + 1], lz = lat[3 * k + 2];
            w.px = p.x - (float)lx * dx0;
            w.py = p.y - (float)ly * dy0;
            w.pz = p.z - (float)lz * dz0;
            w.self = lx == 0 && ly == 0 && lz == 0;
        }
        if (!COHERENT) {
            while (w.node < nrec)
                w.node = bh_step<SYM>(SP, theta, cb, sb, w.node, t.cq[w.node], t.aux[w.node], w, ncell, nion, nopen);
        } else {
        // The next record is usually the following one (a leaf, an opened
        // cell): it is loaded while the current one is processed.
        int r = __builtin_amdgcn_readfirstlane(wave_min(w.node));
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        int4 x = make_int4(0, 0, 0, 0);
        if (r < nrec) {
            q = t.cq[r];
            x = t.aux[r];
        }
        while (r < nrec) {  // uniform: ends when every lane is past the last record
            const int rn = r + 1 < nrec ? r + 1 : r;
            const float4 qn = t.cq[rn];
            const int4 xn = t.aux[rn];
            if (w.node == r) w.node = bh_step<SYM>(SP, theta, cb, sb, r, q, x, w, ncell, nion, nopen);
            const int r2 = __builtin_amdgcn_readfirstlane(wave_min(w.node));
            if (r2 == r + 1) {
                q = qn;
                x = xn;
            } else if (r2 < nrec) {
                q = t.cq[r2];
                x = t.aux[r2];
            }
            r = r2;
        }
        }
        tx += (float)((double)w.ax * kc);
        ty += (float)((double)w.ay * kc);
        tz += (float)((double)w.az * kc);
    }
    if (active) {
        ax_out[i - i0] = tx;
        ay_out[i - i0] = ty;
        az_out[i - i0] = tz;
    }
    unsigned long long a = ncell, b = nion, o = nopen;
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        o += __shfl_xor(o, off);
    }
    if ((threadIdx.x & 63) == 0 && (a | b | o)) {
        atomicAdd(&ctr[0], a);
        atomicAdd(&ctr[1], b);
        atomicAdd(&ctr[3], o);
    }
}

template <bool COHERENT>
void rc_walk(int symt, dim3 g, dim3 b, hipStream_t stream, BhWalk w, int nrec, const float4* SP, const int* SI, int ns,
             const int* st, const int* sp, float theta, const int* lat, int nimg, const float* dom, int i0, double* ax,
             double* ay, double* az, unsigned long long* ctr) {
    if (symt == SHP_SYM_NONE)
        hipLaunchKernelGGL((k_bh_walk<SHP_SYM_NONE, COHERENT>), g, b, 0, stream, w, nrec, SP, SI, ns, st, sp, theta,
                           lat, nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az, ctr);
    else if (symt == SHP_SYM_ROTATIONAL)
        hipLaunchKernelGGL((k_bh_walk<SHP_SYM_ROTATIONAL, COHERENT>), g, b, 0, stream, w, nrec, SP, SI, ns, st, sp,
                           theta, lat, nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az, ctr);
    else
        hipLaunchKernelGGL((k_bh_walk<SHP_SYM_TRANSLATIONAL, COHERENT>), g, b, 0, stream, w, nrec, SP, SI, ns, st, sp,
                           theta, lat, nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az, ctr);
}

__global__ void k_bh_zero(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_bh_zero_u64(unsigned long long* __restrict__ p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_bh_interleave(const double* __restrict__ a, int n, double* __restrict__ out3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out3[3 * (size_t)i] = a[i];
        out3[3 * (size_t)i + 1] = a[n + i];
        out3[3 * (size_t)i + 2] = a[2 * (size_t)n + i];
    }
}

// Counters of the last evaluation (device): [0] deepest level, [1] bucket
// leaves, [2] accepted cells, [3] ion terms, [4] ions placed in the last root
// cell because no cell matched, [5] opened cells.
constexpr int kBhCtr = 6;

// Evaluates targets (status 1) in [i0, i1) from sources (status 1/2) in the
// same range; outputs ax/ay/az indexed by i - i0 (zeroed first).
int bh_eval(shp_handle h, const float4* P, const int* st, const int* sp, int i0, int i1, float theta,
            const shp_symmetry* sym, double* ax, double* ay, double* az) {
    hipStream_t stream = h->stream;
    const int m = i1 - i0;
    h->bh_have_stats = false;
    hipLaunchKernelGGL(k_bh_zero, dim3(bdiv(m, 256)), dim3(256), 0, stream, ax, (size_t)m);
    hipLaunchKernelGGL(k_bh_zero, dim3(bdiv(m, 256)), dim3(256), 0, stream, ay, (size_t)m);
    hipLaunchKernelGGL(k_bh_zero, dim3(bdiv(m, 256)), dim3(256), 0, stream, az, (size_t)m);
    if (m == 0) return SHP_OK;
    const int symt = sym ? sym->type : SHP_SYM_NONE;
    SHP_REQUIRE(symt == SHP_SYM_NONE || symt == SHP_SYM_ROTATIONAL || symt == SHP_SYM_TRANSLATIONAL,
                "shp_bh: unknown symmetry type %d", symt);
    if (symt == SHP_SYM_ROTATIONAL && sym->nsec != 4) {
        set_error("shp_bh: the Barnes-Hut rotational variant has four quadrants (nsec=%d)", sym->nsec);
        return SHP_ERR_UNSUPPORTED;
    }
    SHP_REQUIRE(symt != SHP_SYM_TRANSLATIONAL || (sym->nimg >= 1 && sym->lattice),
                "shp_bh: translational symmetry needs a lattice");
    // sources and their float bounds
    SHP_WS(flags, char, h, WS_FLAGS, m);
    SHP_WS(misc, int, h, WS_COUNTS, 128);  // ints 0..63: tree counters, then the header and counters
    BhHead* dhead = (BhHead*)(misc + 64);
    int* nsel = misc + 72;
    int* nruns = misc + 73;
    unsigned long long* ctr = (unsigned long long*)(misc + 80);
    SHP_WS(src, int, h, WS_BH_SRC, m);
    hipLaunchKernelGGL(k_bh_flags, dim3(bdiv(m, 256)), dim3(256), 0, stream, st, i0, m, flags);
    hipcub::CountingInputIterator<int> it(i0);
    size_t tb = 0;
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, flags, src, nsel, m, stream));
    SHP_WS(cub, char, h, WS_CUB, tb);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, tb, it, flags, src, nsel, m, stream));
    const int nbp = std::min(bdiv(m, kBhThreads), 1024);
    SHP_WS(part, float, h, WS_BBOX, 6 * (size_t)nbp);
    hipLaunchKernelGGL(k_bh_bbox_partial, dim3(nbp), dim3(kBhThreads), 0, stream, P, src, nsel, part);
    hipLaunchKernelGGL(k_bh_bbox_final, dim3(1), dim3(64), 0, stream, part, nbp, nsel, dhead);
    hipLaunchKernelGGL(k_bh_zero_u64, dim3(1), dim3(64), 0, stream, ctr, kBhCtr);
    BhHead head;
    float4 seed4;
    SHP_HIP_CHECK(hipMemcpyAsync(&head, dhead, sizeof(head), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(&seed4, P + i0, sizeof(seed4), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    const int ns = head.ns;
    if (ns == 0) return SHP_OK;
    // the grid of root cells (host), per-level dimensions
    BhLayout L;
    const float seed[3] = {seed4.x, seed4.y, seed4.z};
    bh_layout_host(head.mn, head.mx, seed, ns, h->bh_opt.domain_toggle,
                   h->bh_opt.force_box_dims ? h->bh_opt.box_dims : nullptr, &L);
    const long nroot = (long)L.nd[0] * L.nd[1] * L.nd[2];
    SHP_REQUIRE(nroot >= 1 && nroot <= kBhMaxRoots, "shp_bh: %ld root cells (limit %d)", nroot, kBhMaxRoots);
    BhLevels lv;
    bh_levels_host(L, &lv);
    // keys: root cell, then 42 octant digits in two words
    SHP_WS(rk, unsigned, h, WS_TGT, 3 * (size_t)ns);  // rk, rk2, rkS
    SHP_WS(kk, uint64_t, h, WS_KEYS, 4 * (size_t)ns); // kb, kc, scratch keys, sorted keys
    SHP_WS(sk, uint64_t, h, WS_KEYS2, 2 * (size_t)ns);  // KB, KC in final order
    SHP_WS(vals, int, h, WS_VALS2, 3 * (size_t)ns);
    unsigned *rk2 = rk + ns, *rkS = rk + 2 * (size_t)ns;
    uint64_t *kb = kk, *kc = kk + ns, *kt = kk + 2 * (size_t)ns, *ks = kk + 3 * (size_t)ns;
    int *va = vals, *vb = vals + ns, *vc = vals + 2 * (size_t)ns;
    const dim3 gs(bdiv(ns, 256)), bs(256);
    hipLaunchKernelGGL(k_bh_keys, gs, bs, 0, stream, P, src, ns, L, lv, rk, kb, kc, va, ctr + 4);
    // LSD: levels 22-42, then levels 1-21, then the root cell (stable passes)
    size_t t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, kc, ks, va, vb, ns, 0, 64, stream));
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t2, (const unsigned*)rk2, rkS, vb, vc, ns, 0,
                                                     32, stream));
    SHP_HIP_CHECK(hipcub::DeviceRunLengthEncode::Encode(nullptr, t3, rkS, rk2, vb, nruns, ns, stream));
    t4 = std::max(std::max(t1, t2), t3);
    SHP_WS(cub2, char, h, WS_CUB, t4);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub2, t1, kc, ks, va, vb, ns, 0, 64, stream));
    hipLaunchKernelGGL(k_bh_permute<uint64_t>, gs, bs, 0, stream, kb, vb, ns, kt);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub2, t1, kt, ks, vb, va, ns, 0, 64, stream));
    const int* order = va;
    if (nroot > 1) {
        hipLaunchKernelGGL(k_bh_permute<unsigned>, gs, bs, 0, stream, rk, va, ns, rk2);
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub2, t2, (const unsigned*)rk2, rkS, va, vc, ns, 0,
                                                         32, stream));
        order = vc;
    } else {
        SHP_HIP_CHECK(hipMemcpyAsync(rkS, rk, sizeof(unsigned) * ns, hipMemcpyDeviceToDevice, stream));
    }
    uint64_t *KB = sk, *KC = sk + ns;
    hipLaunchKernelGGL(k_bh_permute<uint64_t>, gs, bs, 0, stream, kb, order, ns, KB);
    hipLaunchKernelGGL(k_bh_permute<uint64_t>, gs, bs, 0, stream, kc, order, ns, KC);
    SHP_WS(SP, float4, h, WS_BH_SP, ns);
    SHP_WS(SI, int, h, WS_BH_SI, ns);
    hipLaunchKernelGGL(k_bh_gather, gs, bs, 0, stream, P, src, order, ns, SP, SI);
    // non-empty root cells: run-length encoding of the sorted root indices
    unsigned* uniq = rk2;
    int* rcnt = (order == vc) ? va : vb;
    SHP_HIP_CHECK(hipcub::DeviceRunLengthEncode::Encode(cub2, t3, rkS, uniq, rcnt, nruns, ns, stream));
    // tree, level by level; retried with room for every level if the first
    // allocation overflows (deeply clustered ions)
    const long nr0 = std::min<long>(nroot, ns);
    for (int attempt = 0; attempt < 2; attempt++) {
        const size_t cap = attempt == 0 ? 4 * (size_t)ns + (size_t)nr0 + 64
                                        : (size_t)(kBhDepth + 1) * ns + (size_t)nr0 + 64;
        SHP_REQUIRE(cap < ((size_t)1 << 31), "shp_bh: %d ions need too many tree nodes", ns);
        const size_t per = 5 * sizeof(int) + 2 * sizeof(float4) + sizeof(float2);
        SHP_WS(nodes, char, h, WS_BH_NODES, cap * per);
        BhNodes t;
        t.geo = (float4*)nodes;
        t.cq = t.geo + cap;
        t.du = (float2*)(t.cq + cap);
        t.start = (int*)(t.du + cap);
        t.count = t.start + cap;
        t.first = t.count + cap;
        t.nch = t.first + cap;
        t.next = t.nch + cap;
        hipLaunchKernelGGL(k_bh_init, dim3(bdiv(nr0, 256)), dim3(256), 0, stream, t, uniq, rcnt, nruns, rkS,
                           ns, L, lv.d[0].w, misc);
        long maxlev = nr0;  // bound on the nodes of a level: min(nroot 8^lev, ns)
        for (int lev = 0; lev < kBhDepth; lev++) {
            const int L1 = lev + 1;
            const uint64_t* key = L1 <= kBhWord ? KB : KC;
            const int shift = 3 * (kBhWord - 1 - (L1 - 1) % kBhWord);
            hipLaunchKernelGGL(k_bh_split, dim3(bdiv(maxlev, 256)), dim3(256), 0, stream, t, lev, key, shift,
                               lv.d[L1], (int)cap, misc);
            hipLaunchKernelGGL(k_bh_mark, dim3(1), dim3(64), 0, stream, lev, (int)cap, misc);
            maxlev = std::min<long>(maxlev * 8, ns);
        }
        int hm[2] = {0, 0};
        SHP_HIP_CHECK(hipMemcpyAsync(hm, misc, 2 * sizeof(int), hipMemcpyDeviceToHost, stream));
        SHP_HIP_CHECK(hipStreamSynchronize(stream));
        if (hm[1] && attempt == 0) continue;
        SHP_REQUIRE(!hm[1], "shp_bh: tree node capacity exceeded");
        const int nodes_total = hm[0];
        long bound = ns;
        for (int lev = kBhDepth; lev >= 0; lev--) {
            bound = lev == 0 ? nr0 : ns;
            hipLaunchKernelGGL(k_bh_props, dim3(bdiv(bound, 256)), dim3(256), 0, stream, t, lev, lv.d[lev + 1],
                               SP, misc);
        }
        hipLaunchKernelGGL(k_bh_tree_stats, dim3(bdiv(ns, 256)), dim3(256), 0, stream, t, misc, ctr);
        // walk records in preorder: subtree sizes bottom-up, the root cells'
        // positions by a scan of their sizes, the others top-down
        SHP_WS(wrec, char, h, WS_BH_WALK, cap * (sizeof(float4) + sizeof(int4) + 2 * sizeof(int)));
        BhWalk w;
        w.cq = (float4*)wrec;
        w.aux = (int4*)(w.cq + cap);
        int* size = (int*)(w.aux + cap);
        int* pre = size + cap;
        for (int lev = kBhDepth; lev >= 0; lev--)
            hipLaunchKernelGGL(k_bh_size, dim3(bdiv(lev == 0 ? nr0 : ns, 256)), dim3(256), 0, stream, t, lev, size,
                               misc);
        size_t t4 = 0;
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, size, pre, (int)nr0, stream));
        SHP_WS(cub4, char, h, WS_CUB, t4);
        SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub4, t4, size, pre, (int)nr0, stream));
        for (int lev = 0; lev < kBhDepth; lev++)
            hipLaunchKernelGGL(k_bh_pre, dim3(bdiv(lev == 0 ? nr0 : ns, 256)), dim3(256), 0, stream, t, lev, size, pre,
                               misc);
        hipLaunchKernelGGL(k_bh_pack, dim3(bdiv(nodes_total, 256)), dim3(256), 0, stream, t, w, size, pre, misc);
        const int* lat = nullptr;
        int nimg = 1;
        float dom[3] = {0.0f, 0.0f, 0.0f};
        if (symt == SHP_SYM_TRANSLATIONAL) {
            nimg = sym->nimg;
            SHP_WS(dl, int, h, WS_IMAGES, 3 * (size_t)nimg);
            SHP_HIP_CHECK(hipMemcpyAsync(dl, sym->lattice, sizeof(int) * 3 * (size_t)nimg,
                                         hipMemcpyHostToDevice, stream));
            lat = dl;
            for (int c = 0; c < 3; c++) dom[c] = (float)sym->domain[c];
        }
        SHP_HIP_CHECK(hipEventRecord(h->ev0, stream));
        const dim3 g(bdiv(ns, kBhThreads)), b(kBhThreads);
        // the coherent walk needs enough waves per SIMD to hide its per-step chain
        const bool coherent = h->bh_opt.walk == 2 || (h->bh_opt.walk == 0 && (long)ns >= kBhCoherentMin);
        if (coherent)
            rc_walk<true>(symt, g, b, stream, w, nodes_total, SP, SI, ns, st, sp, theta, lat, nimg, dom, i0, ax, ay,
                          az, ctr + 2);
        else
            rc_walk<false>(symt, g, b, stream, w, nodes_total, SP, SI, ns, st, sp, theta, lat, nimg, dom, i0, ax, ay,
                           az, ctr + 2);
        SHP_HIP_CHECK(hipGetLastError());
        SHP_HIP_CHECK(hipEventRecord(h->ev1, stream));
        h->timed = true;
        h->bh_ctr = ctr;
        h->bh_have_stats = true;
        h->bh_host_stats[0] = (double)nroot;
        h->bh_host_stats[1] = (double)nodes_total;
        for (int c = 0; c < 3; c++) {
            h->bh_host_stats[2 + c] = (double)L.nd[c];
            h->bh_host_stats[5 + c] = (double)L.roots[c];
            h->bh_host_stats[8 + c] = (double)L.dims[c];
        }
        if (symt == SHP_SYM_TRANSLATIONAL) SHP_HIP_CHECK(hipStreamSynchronize(stream));  // lattice copy
        break;
    }
    return SHP_OK;
}

}  // namespace
}  // namespace shp

using namespace shp;

extern "C" {

int shp_bh_set_options(shp_handle h, const shp_bh_options* o) {
    SHP_REQUIRE(h && o, "shp_bh_set_options: null argument");
    SHP_REQUIRE(o->domain_toggle == 0 || o->domain_toggle == 1,
                "shp_bh_set_options: domain_toggle %d (0 cubic, 1 rectangular prism)", o->domain_toggle);
    if (o->force_box_dims)
        SHP_REQUIRE(o->box_dims[0] >= 1 && o->box_dims[1] >= 1 && o->box_dims[2] >= 1,
                    "shp_bh_set_options: box_dims (%d, %d, %d) must be >= 1", o->box_dims[0], o->box_dims[1],
                    o->box_dims[2]);
    if (o->subcell_block != 2) {
        set_error("shp_bh_set_options: subcell_block %d is not supported (2 x 2 x 2 subcells only)",
                  o->subcell_block);
        return SHP_ERR_UNSUPPORTED;
    }
    if (o->outlier_toggle != 0) {
        set_error("shp_bh_set_options: the outlier toggle is not supported (reference default: off)");
        return SHP_ERR_UNSUPPORTED;
    }
    SHP_REQUIRE(o->walk >= 0 && o->walk <= 2, "shp_bh_set_options: walk %d (0, 1 or 2)", o->walk);
    h->bh_opt = *o;
    return SHP_OK;
}

int shp_bh_stats(shp_handle h, double* out) {
    SHP_REQUIRE(h && out, "shp_bh_stats: null argument");
    SHP_REQUIRE(h->bh_have_stats, "shp_bh_stats: no Barnes-Hut evaluation with sources yet");
    SHP_HIP_CHECK(hipSetDevice(h->device));
    unsigned long long c[kBhCtr];
    SHP_HIP_CHECK(hipMemcpyAsync(c, h->bh_ctr, sizeof(c), hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    for (int k = 0; k < SHP_BH_NSTATS; k++) out[k] = 0.0;
    out[0] = h->bh_host_stats[0];
    out[1] = h->bh_host_stats[1];
    out[2] = (double)c[0];
    out[3] = (double)c[1];
    out[4] = (double)c[2];
    out[5] = (double)c[3];
    out[6] = (double)c[4];
    for (int k = 0; k < 9; k++) out[7 + k] = h->bh_host_stats[2 + k];
    out[16] = (double)c[5];
    return SHP_OK;
}

int shp_bh_dev(shp_handle h, const shp_particles* p, float theta, const shp_symmetry* sym, double* ax,
               double* ay, double* az) {
    SHP_REQUIRE(h && p && ax && ay && az, "shp_bh_dev: null argument");
    SHP_REQUIRE(p->n >= 0, "shp_bh_dev: n=%d", p->n);
    SHP_REQUIRE(p->n == 0 || (p->x && p->y && p->z && p->q && p->status && p->species),
                "shp_bh_dev: null particle array");
    SHP_REQUIRE(theta >= 0.0f, "shp_bh_dev: theta %g < 0", (double)theta);
    SHP_HIP_CHECK(hipSetDevice(h->device));
    const int n = p->n;
    if (n == 0) return SHP_OK;
    SHP_WS(P, float4, h, WS_HOST_IN, n);
    hipLaunchKernelGGL(k_bh_from_f64, dim3(bdiv(n, 256)), dim3(256), 0, h->stream, p->x, p->y, p->z, p->q,
                       n, P);
    return bh_eval(h, P, p->status, p->species, 0, n, theta, sym, ax, ay, az);
}

int shp_bh_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n, int i0,
                int i1, float theta, const shp_symmetry* sym, double* out3) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4 && out3, "shp_bh_host: null argument");
    SHP_REQUIRE(n >= 0 && 0 <= i0 && i0 <= i1 && i1 <= n, "shp_bh_host: bad range n=%d [%d,%d)", n, i0, i1);
    SHP_REQUIRE(theta >= 0.0f, "shp_bh_host: theta %g < 0", (double)theta);
    const int m = i1 - i0;
    if (m == 0) return SHP_OK;
    SHP_HIP_CHECK(hipSetDevice(h->device));
    SHP_WS(dpos, float4, h, WS_HOST_IN, n);
    SHP_WS(dvel, float4, h, WS_HOST_IN2, n);
    SHP_WS(dacc, float4, h, WS_HOST_IN3, n);
    SHP_WS(ints, int, h, WS_WIDE_I32, 2 * (size_t)n);
    SHP_WS(dout, double, h, WS_HOST_OUT, 6 * (size_t)m);
    SHP_HIP_CHECK(hipMemcpyAsync(dpos, pos4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dvel, vel4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dacc, acc4, sizeof(float4) * n, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_bh_from_aos, dim3(bdiv(n, 256)), dim3(256), 0, h->stream, dvel, dacc, n, ints,
                       ints + n);
    int rc = bh_eval(h, dpos, ints, ints + n, i0, i1, theta, sym, dout, dout + m, dout + 2 * (size_t)m);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bh_interleave, dim3(bdiv(m, 256)), dim3(256), 0, h->stream, dout, m, dout + 3 * (size_t)m);
    SHP_HIP_CHECK(hipMemcpyAsync(out3, dout + 3 * (size_t)m, sizeof(double) * 3 * (size_t)m,
                                 hipMemcpyDeviceToHost, h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

}  // extern "C"
