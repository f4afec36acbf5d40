// Barnes-Hut force on gfx950: octree built on the GPU from Morton-sorted
// ions, stackless threaded walk, reference (FP32) arithmetic.
//
// Replaces BarnesHutForceCalculation::updateAcceleration and its passes
// (barnes_hut_force_calculation.cpp:574-715: generateTree :258-331,
// locateSubcell :149-216, addToCell :224-253, computeCellProperties
// :337-390, computeForceFromTree :417-450, barnesHutKernelCalculation
// :539-569) with the rotational (barnes_hut_force_calculation_rotational.cpp:28-135)
// and translational (barnes_hut_force_calculation_translational.cpp:190-280)
// variants, and cell.cpp:7-131 for the cell geometry.
//
// Tree.  The reference inserts ions one by one into a pointer tree and splits
// a cell as soon as it holds two ions, so its cells are exactly the octree
// cells holding two or more ions plus their children.  Here every source ion
// gets a 63-bit key by descending 21 levels with the reference's own float
// compares against the cell centres (location + side / 2, children at
// location + bit * side / 2), the keys are radix-sorted, and the tree is
// built level by level: a node is a run of sorted keys with a common prefix;
// a node of two or more ions is split into its non-empty children (8 binary
// searches), which are allocated contiguously in octant order with the
// threaded successor `next` (next sibling, or the parent's successor).  Empty
// children of the reference carry no charge and are skipped by its walk, so
// omitting them changes no sum.  A node at depth 21 is not split (bucket).
//
// Cell properties are formed bottom-up in the reference's order (children in
// octant order): net charge, sum |q|, centre of charge = sum centre |net q| /
// sum |q| (geometric centre when the sum is zero), delta = |centre of charge
// - geometric centre|.  The walk is the reference's: leaves interact
// directly, an inner cell is accepted iff theta (d - delta) > side with d the
// softened float distance, otherwise opened; the successor pointer makes the
// walk stackless.  All float arithmetic follows the reference operation by
// operation (no contraction), so with the same root the result equals the
// reference algorithm's.
//
// Root: one cube around the FP64 bounding box of the sources (side = largest
// extent x (1 + 2^-20), centred), rounded to float.  The reference instead
// lays a grid of root cells from DomainBoundsFunction::computeDomainBounds.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "shp_internal.h"

#pragma clang fp contract(off)

namespace shp {
namespace {

constexpr int kBhDepth = 21;
constexpr int kBhThreads = 256;

inline int bdiv(long a, long b) { return (int)((a + b - 1) / b); }

struct BhNodes {
    int* start;
    int* count;
    int* first;   // first child (-1: leaf)
    int* nch;
    int* next;    // threaded successor (-1: end of walk)
    float4* geo;  // location (corner) xyz, side
    float4* cq;   // centre of charge xyz, net charge
    float2* du;   // delta, sum |q|
};

// What the walk reads of a node, in one 32-byte record: the centre of charge
// and net charge, then (first child, successor, and for a cell delta and side
// as float bits, for a leaf the ion count and first ion).
struct BhWalk {
    float4* cq;
    int4* aux;
};

__global__ void k_bh_pack(BhNodes t, BhWalk w, const int* __restrict__ misc) {
    const int node = blockIdx.x * blockDim.x + threadIdx.x;
    if (node >= misc[0]) return;
    const int f = t.first[node];
    w.cq[node] = t.cq[node];
    w.aux[node] = f < 0 ? make_int4(f, t.next[node], t.count[node], t.start[node])
                        : make_int4(f, t.next[node], __float_as_int(t.du[node].x), __float_as_int(t.geo[node].w));
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

// ------------------------------------------------------------------ root ----

__global__ __launch_bounds__(kBhThreads) void k_bh_bbox_partial(const float4* __restrict__ P,
                                                                const int* __restrict__ src, int ns,
                                                                double* __restrict__ part) {
    __shared__ double red[6][kBhThreads];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < ns; k += gridDim.x * blockDim.x) {
        const float4 p = P[src[k]];
        const double v[3] = {(double)p.x, (double)p.y, (double)p.z};
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
    for (int s = kBhThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int c = 0; c < 3; c++) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + s]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// Root cube (float location, side) from the partial boxes; one wave.
__global__ __launch_bounds__(64) void k_bh_root(const double* __restrict__ part, int nparts,
                                                float4* __restrict__ root) {
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
    if (threadIdx.x == 0) {
        double side = 0.0;
        for (int c = 0; c < 3; c++) side = fmax(side, mx[c] - mn[c]);
        side *= 1.0 + 0x1p-20;
        if (!(side > 0.0)) side = 1.0;
        root[0] = make_float4((float)(0.5 * (mn[0] + mx[0]) - 0.5 * side),
                              (float)(0.5 * (mn[1] + mx[1]) - 0.5 * side),
                              (float)(0.5 * (mn[2] + mx[2]) - 0.5 * side), (float)side);
    }
}

// ------------------------------------------------------------------ keys ----

// 21 octant digits per ion from the reference's float compares (x -> bit 0,
// y -> bit 1, z -> bit 2 of a digit; the first level in the top digit).
__global__ void k_bh_keys(const float4* __restrict__ P, const int* __restrict__ src, int ns,
                          const float4* __restrict__ root, uint64_t* __restrict__ keys,
                          int* __restrict__ vals) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ns) return;
    const float4 p = P[src[k]];
    const float4 r = root[0];
    float lx = r.x, ly = r.y, lz = r.z, side = r.w;
    uint64_t key = 0;
    for (int lev = 0; lev < kBhDepth; lev++) {
        const float h = side / 2.0f;
        const int bx = p.x > lx + h, by = p.y > ly + h, bz = p.z > lz + h;
        lx += (float)bx * h;
        ly += (float)by * h;
        lz += (float)bz * h;
        side = h;
        key = (key << 3) | (uint64_t)(bx | (by << 1) | (bz << 2));
    }
    keys[k] = key;
    vals[k] = k;
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

// misc[0] = nodes allocated, misc[1] = overflow flag, misc[2 + k] = first node
// of level k (k = 0..22).
__global__ void k_bh_init(BhNodes t, int ns, const float4* __restrict__ root, int* __restrict__ misc) {
    if (threadIdx.x != 0) return;
    t.start[0] = 0;
    t.count[0] = ns;
    t.first[0] = -1;
    t.nch[0] = 0;
    t.next[0] = -1;
    t.geo[0] = root[0];
    misc[0] = 1;
    misc[1] = 0;
    misc[2] = 0;
    misc[3] = 1;
}

__global__ void k_bh_split(BhNodes t, int lev, const uint64_t* __restrict__ keys, int cap,
                           int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    const int c = t.count[i];
    if (c < 2) return;
    const int s = t.start[i], e = s + c;
    const int shift = 3 * (kBhDepth - 1 - lev);  // digit of level lev + 1
    int b[9];
    b[0] = s;
    b[8] = e;
    for (int d = 1; d < 8; d++) {  // first key whose digit is >= d
        int lo = b[d - 1], hi = e;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if ((int)((keys[mid] >> shift) & 7) < d) lo = mid + 1;
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
    const float h = g.w / 2.0f;
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
        t.geo[k] = make_float4(g.x + (float)(d & 1) * h, g.y + (float)((d >> 1) & 1) * h,
                               g.z + (float)((d >> 2) & 1) * h, h);
        j++;
    }
    t.first[i] = base;
    t.nch[i] = ne;
}

// Level lev + 1 ends where the allocation stands after splitting level lev.
__global__ void k_bh_mark(int lev, int cap, int* __restrict__ misc) {
    if (threadIdx.x == 0) misc[2 + lev + 2] = min(misc[0], cap);
}

__global__ void k_bh_props(BhNodes t, int lev, const float4* __restrict__ SP,
                           const int* __restrict__ misc) {
    const int i = misc[2 + lev] + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= misc[2 + lev + 1]) return;
    const float4 g = t.geo[i];
    const float h = g.w / 2.0f;
    const float gx = g.x + h, gy = g.y + h, gz = g.z + h;
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

template <int SYM>
__global__ __launch_bounds__(kBhThreads) void k_bh_walk(
    BhWalk t, const float4* __restrict__ SP, const int* __restrict__ SI, int ns,
    const int* __restrict__ st, const int* __restrict__ sp, float theta, const int* __restrict__ lat,
    int nimg, float dx0, float dy0, float dz0, double K, int i0, double* __restrict__ ax_out,
    double* __restrict__ ay_out, double* __restrict__ az_out) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int i = SI[s];
    if (st[i] != 1) return;
    const float4 p = SP[s];
    const double kc = K * (double)p.w / species_mass(sp[i]);
    const int nrep = SYM == SHP_SYM_ROTATIONAL ? 4 : (SYM == SHP_SYM_TRANSLATIONAL ? nimg : 1);
    float tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (int k = 0; k < nrep; k++) {
        float px = p.x, py = p.y, pz = p.z;
        int cb = 1, sb = 0;
        bool self = true;
        if (SYM == SHP_SYM_ROTATIONAL) {
            const int cf = bh_qrf(5 - k), sf = bh_qrf(4 - k);
            cb = bh_qrf(k + 1);
            sb = bh_qrf(k);
            px = (float)cf * p.x - (float)sf * p.y;
            py = (float)sf * p.x + (float)cf * p.y;
            self = k == 0;
        } else if (SYM == SHP_SYM_TRANSLATIONAL) {
            const int lx = lat[3 * k], ly = lat[3 * k + 1], lz = lat[3 * k + 2];
            px = p.x - (float)lx * dx0;
            py = p.y - (float)ly * dy0;
            pz = p.z - (float)lz * dz0;
            self = lx == 0 && ly == 0 && lz == 0;
        }
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        int node = 0;
        while (node >= 0) {
            const float4 q = t.cq[node];
            const int4 x = t.aux[node];
            const int f = x.x;
            if (f < 0) {
                const int c = x.z, j0 = x.w;
                if (c == 1) {
                    if (!(self && j0 == s)) {
                        float rx, ry, rz;
                        const float d = bh_dist(px, py, pz, q.x, q.y, q.z, rx, ry, rz);
                        bh_add<SYM>(q.w / (d * d * d), rx, ry, rz, cb, sb, ax, ay, az);
                    }
                } else {
                    for (int j = j0; j < j0 + c; j++) {
                        if (self && j == s) continue;
                        const float4 q = SP[j];
                        float rx, ry, rz;
                        const float d = bh_dist(px, py, pz, q.x, q.y, q.z, rx, ry, rz);
                        bh_add<SYM>(q.w / (d * d * d), rx, ry, rz, cb, sb, ax, ay, az);
                    }
                }
                node = x.y;
            } else {
                float rx, ry, rz;
                const float d = bh_dist(px, py, pz, q.x, q.y, q.z, rx, ry, rz);
                if (theta * (d - __int_as_float(x.z)) > __int_as_float(x.w)) {
                    bh_add<SYM>(q.w / (d * d * d), rx, ry, rz, cb, sb, ax, ay, az);
                    node = x.y;
                } else {
                    node = f;
                }
            }
        }
        tx += (float)((double)ax * kc);
        ty += (float)((double)ay * kc);
        tz += (float)((double)az * kc);
    }
    ax_out[i - i0] = tx;
    ay_out[i - i0] = ty;
    az_out[i - i0] = tz;
}

__global__ void k_bh_zero(double* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}

__global__ void k_bh_interleave(const double* __restrict__ a, int n, double* __restrict__ out3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out3[3 * (size_t)i] = a[i];
        out3[3 * (size_t)i + 1] = a[n + i];
        out3[3 * (size_t)i + 2] = a[2 * (size_t)n + i];
    }
}

// Evaluates targets (status 1) in [i0, i1) from sources (status 1/2) in the
// same range; outputs ax/ay/az indexed by i - i0 (zeroed first).
int bh_eval(shp_handle h, const float4* P, const int* st, const int* sp, int i0, int i1, float theta,
            const shp_symmetry* sym, double* ax, double* ay, double* az) {
    hipStream_t stream = h->stream;
    const int m = i1 - i0;
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
    // sources
    SHP_WS(flags, char, h, WS_FLAGS, m);
    SHP_WS(misc, int, h, WS_COUNTS, 32);
    SHP_WS(src, int, h, WS_BH_SRC, m);
    hipLaunchKernelGGL(k_bh_flags, dim3(bdiv(m, 256)), dim3(256), 0, stream, st, i0, m, flags);
    hipcub::CountingInputIterator<int> it(i0);
    size_t tb = 0;
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, flags, src, misc + 28, m, stream));
    SHP_WS(cub, char, h, WS_CUB, tb);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, tb, it, flags, src, misc + 28, m, stream));
    int ns = 0;
    SHP_HIP_CHECK(hipMemcpyAsync(&ns, misc + 28, sizeof(int), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    if (ns == 0) return SHP_OK;
    // root and keys
    const int nbp = std::min(bdiv(ns, kBhThreads), 1024);
    SHP_WS(part, double, h, WS_BBOX, 6 * (size_t)nbp);
    SHP_WS(root, float4, h, WS_BOX, 1);
    hipLaunchKernelGGL(k_bh_bbox_partial, dim3(nbp), dim3(kBhThreads), 0, stream, P, src, ns, part);
    hipLaunchKernelGGL(k_bh_root, dim3(1), dim3(64), 0, stream, part, nbp, root);
    SHP_WS(keys, uint64_t, h, WS_KEYS, 2 * (size_t)ns);
    SHP_WS(vals, int, h, WS_VALS2, 2 * (size_t)ns);
    hipLaunchKernelGGL(k_bh_keys, dim3(bdiv(ns, 256)), dim3(256), 0, stream, P, src, ns, root, keys, vals);
    size_t ts = 0;
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, ts, keys, keys + ns, vals, vals + ns, ns, 0,
                                                     3 * kBhDepth, stream));
    SHP_WS(cub2, char, h, WS_CUB, ts);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub2, ts, keys, keys + ns, vals, vals + ns, ns, 0,
                                                     3 * kBhDepth, stream));
    const uint64_t* skeys = keys + ns;
    SHP_WS(SP, float4, h, WS_BH_SP, ns);
    SHP_WS(SI, int, h, WS_BH_SI, ns);
    hipLaunchKernelGGL(k_bh_gather, dim3(bdiv(ns, 256)), dim3(256), 0, stream, P, src, vals + ns, ns, SP, SI);
    // tree, level by level; retried with room for every level if the first
    // allocation overflows (deeply clustered ions)
    for (int attempt = 0; attempt < 2; attempt++) {
        const size_t cap = attempt == 0 ? 4 * (size_t)ns + 64 : (size_t)(kBhDepth + 1) * ns + 64;
        SHP_REQUIRE(cap < ((size_t)1 << 31), "shp_bh: %d ions need too many tree nodes", ns);
        const size_t per = 5 * sizeof(int) + 2 * sizeof(float4) + sizeof(float2);
        SHP_WS(nodes, char, h, WS_BH_NODES, cap * per);
        BhNodes t;
        t.start = (int*)nodes;
        t.count = t.start + cap;
        t.first = t.count + cap;
        t.nch = t.first + cap;
        t.next = t.nch + cap;
        t.geo = (float4*)(t.next + cap);
        t.cq = t.geo + cap;
        t.du = (float2*)(t.cq + cap);
        hipLaunchKernelGGL(k_bh_init, dim3(1), dim3(64), 0, stream, t, ns, root, misc);
        long maxlev = 1;  // bound on the nodes of a level: min(8^lev, ns)
        for (int lev = 0; lev < kBhDepth; lev++) {
            hipLaunchKernelGGL(k_bh_split, dim3(bdiv(maxlev, 256)), dim3(256), 0, stream, t, lev, skeys,
                               (int)cap, misc);
            hipLaunchKernelGGL(k_bh_mark, dim3(1), dim3(64), 0, stream, lev, (int)cap, misc);
            maxlev = std::min<long>(maxlev * 8, ns);
        }
        int ovf = 0;
        SHP_HIP_CHECK(hipMemcpyAsync(&ovf, misc + 1, sizeof(int), hipMemcpyDeviceToHost, stream));
        SHP_HIP_CHECK(hipStreamSynchronize(stream));
        if (ovf && attempt == 0) continue;
        SHP_REQUIRE(!ovf, "shp_bh: tree node capacity exceeded");
        for (int lev = kBhDepth; lev >= 0; lev--) {
            const long bound = std::min<long>(lev < 8 ? (1L << (3 * lev)) : (long)ns, (long)ns);
            hipLaunchKernelGGL(k_bh_props, dim3(bdiv(bound, 256)), dim3(256), 0, stream, t, lev, SP, misc);
        }
        SHP_WS(wrec, char, h, WS_BH_WALK, cap * (sizeof(float4) + sizeof(int4)));
        BhWalk w;
        w.cq = (float4*)wrec;
        w.aux = (int4*)(w.cq + cap);
        hipLaunchKernelGGL(k_bh_pack, dim3(bdiv((long)cap, 256)), dim3(256), 0, stream, t, w, misc);
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
        if (symt == SHP_SYM_NONE)
            hipLaunchKernelGGL(k_bh_walk<SHP_SYM_NONE>, g, b, 0, stream, w, SP, SI, ns, st, sp, theta, lat,
                               nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az);
        else if (symt == SHP_SYM_ROTATIONAL)
            hipLaunchKernelGGL(k_bh_walk<SHP_SYM_ROTATIONAL>, g, b, 0, stream, w, SP, SI, ns, st, sp, theta,
                               lat, nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az);
        else
            hipLaunchKernelGGL(k_bh_walk<SHP_SYM_TRANSLATIONAL>, g, b, 0, stream, w, SP, SI, ns, st, sp,
                               theta, lat, nimg, dom[0], dom[1], dom[2], k_const_host(), i0, ax, ay, az);
        SHP_HIP_CHECK(hipGetLastError());
        SHP_HIP_CHECK(hipEventRecord(h->ev1, stream));
        h->timed = true;
        if (symt == SHP_SYM_TRANSLATIONAL) SHP_HIP_CHECK(hipStreamSynchronize(stream));  // lattice copy
        break;
    }
    return SHP_OK;
}

}  // namespace
}  // namespace shp

using namespace shp;

extern "C" {

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
