// The reference's own FMM operators on the GPU: spherical-harmonic expansions
// of 1/r, translated by a rotation, a shift along z and the rotation back
// (MultipoleMethodForceCalculation, fmm_cpukernel.cpp:10-1317, tables of
// MultipoleMethodConstants::precalc, multipole_method_constants.cpp:40-226),
// over the reference's domain grid and leaf placement, in FP64.
//
// MI355X mapping.  Every translation of the reference is a fixed real-linear
// map of the 2 p(p+1)/2 real components of an expansion: M2M and L2L per child
// octant (level independent), M2L per offset and level (the softened centre
// distance depends on the box size).  The engine builds each map once per call
// as a dense matrix by applying the reference's rotate / shift / rotate
// sequence to unit vectors (one thread per column), and then runs the M2L of
// a level as batched dense products on the FP64 matrix cores: a wave takes 16
// target boxes of one child octant (they share the offset list), gathers the
// 16 source expansions of each offset as the B operand and accumulates
// T(offset) x M with v_mfma_f64_16x16x4f64, the operator pre-swizzled so each
// A fragment is one coalesced 512-byte load.  P2M and L2P follow the
// reference's Legendre recurrences per particle; P2P is the reference's direct
// sum over the 27 neighbour leaves.
//
// Tree.  The reference lays a grid of cubic domains of side rootBoxSize
// (computeDomainBounds, setDomainSize multipole_method_force_calculation.cpp:
// 648-832), gives every domain an octree of depth maxLevel (setOptimumLevel
// :835-881 or setMaxLevel) and places each particle with float arithmetic
// (morton :976-1057).  Below the domain level its P2P and M2L lists
// (getInteractionList :1685-2446) are those of one octree over a cube of side
// 2^K rootBoxSize (2^K >= the largest domain count), which is the tree used
// here at every level; for one domain (K = 0) it is the reference's tree.
// Above the domain level the reference groups domains into ternary
// supercells; here they are octree levels of the enclosing cube.
//
// Rotational symmetry (updateAccelerationRotational, multipole_method_force_
// calculation_rotational.cpp:9-331): the ions fill quadrant 0 (x >= 0, y >= 0)
// of one column of domains centred on the z axis (computeDomainBoundsRotations)
// and quadrants 1-3 hold their images rotated by k pi / 2.  Below the domain
// level a cell of the local window in quadrant k stands for the real cell it
// images (getInteractionListRotations :568-1438): its P2P sources are rotated
// in float (direct_rotational), its multipole is multiplied by i^(-m k) in the
// M2L gather (m2l_kernel_rotations, fmm_cpukernel_rotational.cpp:298-496), and
// the M2M into a domain adds the four images of each child (m2m_kernel_rotations
// :107-290), one dense operator per child z.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "shp_internal.h"

namespace shp {
namespace {

constexpr int kSpPmax = 10;     // FMMConstant::maxNumExpansions (constants.h:82)
constexpr int kSpNe2 = 100;     // numExpansion2
constexpr int kSpNe4 = 10000;   // numExpansion4
constexpr int kSpNrot = 1155;   // possibleRotations
constexpr int kSpNrel = 1331;   // numRelativeBox = 11^3
constexpr int kSpNcMax = 55;    // numCoefficients at p = 10
constexpr int kSpOps = 343;     // M2L operator slots: offsets in [-3, 3]^3
constexpr int kSpNoff = 189;    // M2L offsets per child octant
constexpr int kSpP2pTile = 256; // P2P sources staged in LDS per barrier pair
constexpr int kSpMaxDomains = 1 << 22;
constexpr double kPiHi = 3.141592653589793;       // long double pi as a double pair
constexpr double kPiLo = 1.2246467991473532e-16;
constexpr double kPiHalf = 1.5707963267948966;

// ------------------------------------------------------------ host tables --

struct SphHostTables {
    float factorial[4 * kSpNe2];
    float anm[4 * kSpNe2];
    float Anm[kSpNe4];
    double Ynm[4 * kSpNe2];  // the reference's complex Ynm table is real
    double2 Dnm[kSpNrot][kSpPmax][kSpNe2];
    int link[kSpNrel];
};

// a + b c with the product rounded before the sum, as the reference's x86-64
// build computes it (no fused multiply-add): values that are rounded to float
// afterwards (box corners, centres, cart2sph arguments) must not be fused, or
// the float rounding can differ from the reference's.
__host__ __device__ inline float sp_madd_unfused(float a, float b, float c) {
#pragma clang fp contract(off)
    return a + b * c;
}
__host__ __device__ inline double sp_madd_unfused(double a, double b, double c) {
#pragma clang fp contract(off)
    return a + b * c;
}
__host__ __device__ inline double sp_sq3_unfused(double x, double y, double z) {
#pragma clang fp contract(off)
    return x * x + y * y + z * z;
}

// cart2sph_exact (auxiliary_functions.cpp:160-186): float sqrtf / acosf /
// atanf on double arguments, phi + pi formed in long double (PI is long
// double, constants.h:28).
__host__ __device__ inline void sp_cart2sph(double& r, double& th, double& ph, double dx, double dy, double dz) {
    r = (double)(float)sqrt((double)(float)sp_sq3_unfused(dx, dy, dz));
    if (r != 0) {
        th = (double)(float)acos((double)(float)(dz / r));
        if (th == 0 && dz != 0) th = (double)(float)sqrt((double)(float)sp_sq3_unfused(dx, dy, 0.0)) / dz;
    } else {
        th = 0;
    }
    if (dx == 0 && dy == 0) {
        ph = 0;
    } else if (dx == 0) {
        ph = dy > 0 ? kPiHalf : -kPiHalf;
    } else {
        const double a = (double)(float)atan((double)(float)(dy / dx));
        if (dx > 0) {
            ph = a;
        } else {
            const double s = a + kPiHi;  // a + pi rounded once (two-sum of a and pi's leading part)
            const double bb = s - a;
            const double err = (a - (s - bb)) + (kPiHi - bb);
            ph = s + (err + kPiLo);
        }
    }
}

__host__ __device__ inline int sp_gcd(int a, int b) {
    a = a < 0 ? -a : a;
    b = b < 0 ? -b : b;
    while (b != 0) {
        const int t = b;
        b = a % b;
        a = t;
    }
    return a;
}

inline int sp_nmk(int n, int m, int k) { return (4 * n * n * n + 6 * n * n + 5 * n) / 3 + m * (2 * n + 1) + k; }

// MultipoleMethodConstants::precalc: the same recurrences and roundings
// (float storage of factorial, anm and Anm).
void sp_build_tables(SphHostTables* T) {
    const int P = kSpPmax;
    for (int n = 0; n < 2 * P; n++)
        for (int m = -n; m <= n; m++) {
            const int nm = n * n + n + m, am = std::abs(m);
            double fnmm = 1, fnpm = 1, fnma = 1, fnpa = 1;
            for (int i = 1; i <= n - m; i++) fnmm *= i;
            for (int i = 1; i <= n + m; i++) fnpm *= i;
            for (int i = 1; i <= n - am; i++) fnma *= i;
            for (int i = 1; i <= n + am; i++) fnpa *= i;
            T->factorial[nm] = (float)std::sqrt(fnma / fnpa);
            T->anm[nm] = (float)(std::pow(-1.0, n) / std::sqrt(fnmm * fnpm));
        }
    std::memset(T->Anm, 0, sizeof(T->Anm));
    for (int j = 0; j < P; j++)
        for (int k = -j; k <= j; k++) {
            const int jk = j * j + j + k;
            for (int n = std::abs(k); n < P; n++) {
                const int nk = n * n + n + k;
                const int jnk = (j + n) * (j + n) + j + n;
                T->Anm[jk * kSpNe2 + nk] = (float)(std::pow(-1.0, j + k) * T->anm[nk] * T->anm[jk] / T->anm[jnk]);
            }
        }
    double pn = 1;
    for (int m = 0; m < 2 * P; m++) {
        double p = pn;
        T->Ynm[m * m + 2 * m] = T->factorial[m * m + 2 * m] * p;
        T->Ynm[m * m] = T->Ynm[m * m + 2 * m];
        double p1 = p;
        p = (2 * m + 1) * p;
        for (int n = m + 1; n < 2 * P; n++) {
            T->Ynm[n * n + n + m] = T->factorial[n * n + n + m] * p;
            T->Ynm[n * n + n - m] = T->Ynm[n * n + n + m];
            const double p2 = p1;
            p1 = p;
            p = ((2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
        }
        pn = 0;
    }
    std::vector<double> a0(kSpNe4), a1(kSpNe4), D(kSpNe4);
    for (int n = 0; n < P; n++)
        for (int m = 1; m <= n; m++) {
            const double anmd = n * (n + 1) - m * (m - 1);
            for (int k = 1 - m; k < m; k++) {
                const int i = sp_nmk(n, m, k);
                a0[i] = -(m + k) / std::sqrt(anmd);
                a1[i] = std::sqrt(((double)(n * (n + 1) - k * (k + 1))) / (n * (n + 1) - m * (m - 1)));
            }
        }
    int nrot = 0;
    for (int i = 0; i < kSpNrel; i++) {
        // base-11 digits y, x, z (unmorton11)
        const int by = i % 11, bx = (i / 11) % 11, bz = i / 121;
        const int ix = bx - 5, iy = by - 5, iz = bz - 5;
        if (sp_gcd(ix, sp_gcd(iy, iz)) > 1) {
            T->link[i] = -1;
            continue;
        }
        T->link[i] = nrot;
        double rho, alpha, beta;
        sp_cart2sph(rho, alpha, beta, ix, iy, iz);
        // the host keeps the reference's long double phi + pi exactly
        if (ix < 0) beta = (double)((long double)(float)std::atan((double)(float)((double)iy / ix)) +
                                    3.14159265358979323846264338328L);
        if (rho == 0) alpha = 0;
        const double sc = std::sin(alpha) / (1 + std::cos(alpha));
        double2 eb[4 * kSpPmax];
        for (int n = 0; n < 4 * P - 3; n++) {
            const double t = (n - 2 * P + 2) * beta;
            eb[n] = make_double2(std::cos(t), std::sin(t));
        }
        const bool south = (rho == -(double)iz) && (rho > 0);
        for (int n = 0; n < P; n++) {
            D[sp_nmk(n, n, n)] = std::pow(std::cos(alpha * 0.5), 2 * n);
            if (south && n > 0) D[sp_nmk(n, n, n)] = 0;
            for (int k = n; k >= 1 - n; k--) {
                const double ank = ((double)n + k) / (n - k + 1);
                D[sp_nmk(n, n, k - 1)] = std::sqrt(ank) * std::tan(alpha * 0.5) * D[sp_nmk(n, n, k)];
                if (south) D[sp_nmk(n, n, k - 1)] = (k == 1 - n) ? 1 : 0;
            }
            int count = 0;
            for (int m = n; m >= 1; m--)
                for (int k = m - 1; k >= 1 - m; k--) {
                    const int a = sp_nmk(n, m, k);
                    D[sp_nmk(n, m - 1, k)] = a1[a] * D[sp_nmk(n, m, k + 1)] + a0[a] * sc * D[a];
                    if (south) {
                        count++;
                        D[sp_nmk(n, m - 1, k)] = (k == 1 - m) ? std::pow(-1.0, count) : 0;
                    }
                }
        }
        for (int n = 1; n < P; n++)
            for (int m = 0; m <= n; m++) {
                const int base = (4 * n * n * n + 6 * n * n + 5 * n) / 3;
                for (int k = -m; k <= -1; k++) {
                    const int a = sp_nmk(n, m, k);
                    D[a] = std::pow(-1.0, k) * D[a];
                    D[base - k * (2 * n + 1) - m] = std::pow(-1.0, m + k) * D[a];
                }
                for (int k = 0; k <= m; k++) {
                    const int a = sp_nmk(n, m, k), b = sp_nmk(n, k, m);
                    D[b] = std::pow(-1.0, m + k) * D[a];
                    D[base - k * (2 * n + 1) - m] = D[b];
                }
            }
        for (int n = 0; n < P; n++)
            for (int m = 0; m <= n; m++)
                for (int k = -n; k <= n; k++) {
                    const double v = D[sp_nmk(n, m, k)];
                    const double2 e = eb[k + m + 2 * P - 2];
                    T->Dnm[nrot][m][n * (n + 1) + k] = make_double2(e.x * v, e.y * v);
                }
        nrot++;
    }
}

const SphHostTables* sp_host_tables() {
    static SphHostTables* T = nullptr;
    static std::once_flag once;
    std::call_once(once, [] {
        T = new SphHostTables();
        sp_build_tables(T);
    });
    return T;
}

// ---------------------------------------------------------- device tables --

struct SphTab {
    const float* factorial;
    const float* anm;
    const float* Anm;
    const double* Ynm;
    const double2* Dnm;  // [rot][m][nk]
    const int* link;
};

struct SphState {
    void* buf = nullptr;
    SphTab tab{};
};

// rotationMorton on the 11-grid (multipole_method_force_calculation.cpp:1610-1652).
__device__ inline void sp_rot_index(const SphTab& t, int x, int y, int z, int& fwd, int& bwd) {
    const int base = y + z * 121;
    fwd = t.link[base + x * 11];
    bwd = t.link[base + (10 - x) * 11];
}

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// The quadrant map of getInteractionListRotations (multipole_method_force_
// calculation_rotational.cpp:616-634): cell (x, y) of a local window of jd
// cells per axis lies in quadrant k about jd / 2 and images the real cell
// (rx, ry) of quadrant 0 rotated by k pi / 2.
__host__ __device__ inline int sp_quadrant(int x, int y, int jd, int& rx, int& ry) {
    const int half = jd >> 1;
    if (x >= half && y >= half) {
        rx = x;
        ry = y;
        return 0;
    }
    if (x < half && y >= half) {
        rx = y;
        ry = jd - x - 1;
        return 1;
    }
    if (x < half && y < half) {
        rx = jd - x - 1;
        ry = jd - y - 1;
        return 2;
    }
    rx = jd - y - 1;
    ry = x;
    return 3;
}

// C * qrf_complex(-m k) = C i^(-m k) (auxiliary_functions.cpp:83-88): a
// multipole of order m imaged into quadrant k.
__device__ inline double2 sp_phase(double2 c, int m, int k) {
    switch ((-m * k) & 3) {
    case 0: return c;
    case 1: return make_double2(-c.y, c.x);
    case 2: return make_double2(-c.x, -c.y);
    default: return make_double2(c.y, -c.x);
    }
}

// MultipoleMethodForceCalculation::rotation (fmm_cpukernel.cpp:10-32).
__device__ void sp_rotate(const SphTab& t, int p, const double2* C, double2* out, int rot) {
    const double2* Dr = t.Dnm + (size_t)rot * kSpPmax * kSpNe2;
    for (int n = 0; n < p; n++)
        for (int m = 0; m <= n; m++) {
            const double2* Dm = Dr + m * kSpNe2 + n * (n + 1);
            double2 s = make_double2(0, 0);
            for (int k = -n; k <= -1; k++) {
                const double2 c = C[n * (n + 1) / 2 - k];
                const double2 v = cmul(Dm[k], make_double2(c.x, -c.y));
                s.x += v.x;
                s.y += v.y;
            }
            for (int k = 0; k <= n; k++) {
                const double2 v = cmul(Dm[k], C[n * (n + 1) / 2 + k]);
                s.x += v.x;
                s.y += v.y;
            }
            out[n * (n + 1) / 2 + m] = s;
        }
}

// The z shifts of m2m_kernel (fmm_cpukernel.cpp:587-601), m2l_kernel (:878-897)
// and l2l_kernel (:1100-1116), in their arithmetic order.
// z shift of a multipole by rho child sizes, scaled base^-j (octree: sqrt(3) / 2
// and 2, m2m_kernel; ternary: sqrt(|x-1| + |y-1| + |z-1|) and 3,
// m2m_kernel_translations, fmm_cpukernel_translational.cpp:526-543).
__device__ void sp_shift_m2m(const SphTab& t, int p, double rho, double base, const double2* B, double2* A) {
    for (int j = 0; j < p; j++) {
        const double powTwo = pow(base, -j);
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            double2 s = make_double2(0, 0);
            for (int n = 0; n <= j - k; n++) {
                const int jnk = (j - n) * (j - n) + j - n + k;
                const int nm = n * n + n;
                const double c0 = pow(-1.0, n) * (double)t.anm[nm] * (double)t.anm[jnk] / (double)t.anm[jk] * pow(rho, n);
                const double cr = t.Ynm[nm] * c0 * powTwo;
                const double2 b = B[(j - n) * (j - n + 1) / 2 + k];
                const double2 v = cmul(b, make_double2(cr, 0.0 * c0 * powTwo));
                s.x += v.x;
                s.y += v.y;
            }
            A[j * (j + 1) / 2 + k] = s;
        }
    }
}

__device__ void sp_shift_m2l(const SphTab& t, int p, double rho, const double2* A, double2* LA) {
    double rhoj = 1;
    for (int j = 0; j < p; j++) {
        double rhojk = rhoj;
        rhoj *= rho;
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            double2 s = make_double2(0, 0);
            double rhojn = rhojk;
            rhojk *= rho;
            for (int n = k; n < p; n++) {
                rhojn *= rho;
                const int nk = n * n + n + k;
                const int jnk = (j + n) * (j + n) + j + n;
                const double c = (double)t.Anm[jk * kSpNe2 + nk] / rhojn;
                const double2 v = cmul(A[n * (n + 1) / 2 + k], make_double2(t.Ynm[jnk] * c, 0.0 * c));
                s.x += v.x;
                s.y += v.y;
            }
            LA[j * (j + 1) / 2 + k] = s;
        }
    }
}

// z shift of a local expansion by rho child sizes, scaled base^(-n-1) (l2l_kernel;
// l2l_kernel_translations, fmm_cpukernel_translational.cpp:812-826).
__device__ void sp_shift_l2l(const SphTab& t, int p, double rho, double base, const double2* B, double2* A) {
    for (int j = 0; j < p; j++)
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            double2 s = make_double2(0, 0);
            for (int n = j; n < p; n++) {
                const int jnk = (n - j) * (n - j) + n - j;
                const int nk = n * n + n + k;
                const float f = t.anm[jnk] * t.anm[jk] / t.anm[nk];  // float arithmetic in the reference
                const double c = (double)f * pow(rho, n - j);
                const double pw = pow(base, -n - 1);
                const double2 v = cmul(B[n * (n + 1) / 2 + k], make_double2(t.Ynm[jnk] * c * pw, 0.0 * c * pw));
                s.x += v.x;
                s.y += v.y;
            }
            A[j * (j + 1) / 2 + k] = s;
        }
}

// One operator, applied to the unit vector of real component `col`.
// kind 0: M2M of child octant a; 1: L2L into child octant a; 4 / 5: M2M / L2L
// of the ternary child of base-27 digit a; 2: M2L of the
// offset (a, b, c) = target - source at box size h; 3: the rotational M2M of a
// child with z bit a into its domain, the sum over its four quadrant images,
// image k placed at octant (x, y) = (1, 1), (0, 1), (0, 0), (1, 0)
// (fmm_cpukernel_rotational.cpp:148-215).
__device__ void sp_apply_op(const SphTab& t, int p, int kind, int a, int b, int c, double h, int col, double2* X,
                            double2* Y) {
    const int nc = p * (p + 1) / 2;
    for (int i = 0; i < nc; i++) X[i] = make_double2(0, 0);
    if (col & 1) X[col >> 1].y = 1.0;
    else X[col >> 1].x = 1.0;
    int fwd, bwd;
    if (kind == 3) {
        double2 Z[kSpNcMax];
        for (int i = 0; i < nc; i++) Z[i] = make_double2(0, 0);
        const int ci = col >> 1;
        int mc = 0;
        while ((mc + 1) * (mc + 2) / 2 <= ci) mc++;
        mc = ci - mc * (mc + 1) / 2;
        const double2 unit = (col & 1) ? make_double2(0.0, 1.0) : make_double2(1.0, 0.0);
        for (int k = 0; k < 4; k++) {
            const int qx = (k == 0 || k == 3) ? 1 : 0, qy = (k < 2) ? 1 : 0;
            for (int i = 0; i < nc; i++) X[i] = make_double2(0, 0);
            X[ci] = sp_phase(unit, mc, k);
            sp_rot_index(t, 6 - 2 * qx, 6 - 2 * qy, 6 - 2 * a, fwd, bwd);
            sp_rotate(t, p, X, Y, fwd);
            sp_shift_m2m(t, p, sqrt(3.0) / 2, 2.0, Y, X);
            sp_rotate(t, p, X, Y, bwd);
            for (int i = 0; i < nc; i++) {
                Z[i].x += Y[i].x;
                Z[i].y += Y[i].y;
            }
        }
        for (int i = 0; i < nc; i++) Y[i] = Z[i];
        return;
    }
    if (kind == 0) {
        const int cx = (a >> 1) & 1, cy = a & 1, cz = (a >> 2) & 1;
        sp_rot_index(t, 6 - 2 * cx, 6 - 2 * cy, 6 - 2 * cz, fwd, bwd);
        sp_rotate(t, p, X, Y, fwd);
        sp_shift_m2m(t, p, sqrt(3.0) / 2, 2.0, Y, X);
        sp_rotate(t, p, X, Y, bwd);
    } else if (kind == 1) {
        const int cx = (a >> 1) & 1, cy = a & 1, cz = (a >> 2) & 1;
        sp_rot_index(t, 2 * cx + 4, 2 * cy + 4, 2 * cz + 4, fwd, bwd);
        sp_rotate(t, p, X, Y, fwd);
        sp_shift_l2l(t, p, sqrt(3.0) / 2, 2.0, Y, X);
        sp_rotate(t, p, X, Y, bwd);
    } else if (kind == 4 || kind == 5) {
        // ternary child of base-27 digit a = 9 z + 3 x + y (unmorton3 of nfjc / nfic)
        const int cx = (a / 3) % 3, cy = a % 3, cz = a / 9;
        const double rho = sqrt((double)(abs(cx - 1) + abs(cy - 1) + abs(cz - 1)));
        if (kind == 4) sp_rot_index(t, 6 - cx, 6 - cy, 6 - cz, fwd, bwd);
        else sp_rot_index(t, cx + 4, cy + 4, cz + 4, fwd, bwd);
        sp_rotate(t, p, X, Y, fwd);
        if (kind == 4) sp_shift_m2m(t, p, rho, 3.0, Y, X);
        else sp_shift_l2l(t, p, rho, 3.0, Y, X);
        sp_rotate(t, p, X, Y, bwd);
    } else {
        const double fx = a, fy = b, fz = c;
        const double rho = sqrt(fx * fx + fy * fy + fz * fz + kSoftening / (h * h));
        int dx = a, dy = b, dz = c;
        const int g = sp_gcd(dx, sp_gcd(dy, dz));
        if (g == 3 || g == 5) {
            dx = (dx > 0) - (dx < 0);
            dy = (dy > 0) - (dy < 0);
            dz = (dz > 0) - (dz < 0);
        } else if (g > 1) {
            dx /= g;
            dy /= g;
            dz /= g;
        }
        sp_rot_index(t, dx + 5, dy + 5, dz + 5, fwd, bwd);
        sp_rotate(t, p, X, Y, fwd);
        sp_shift_m2l(t, p, rho, Y, X);
        sp_rotate(t, p, X, Y, bwd);
    }
}

// ------------------------------------------------------------- tree keys --

// Morton interleave in the reference's bit order (unmorton,
// multipole_method_functions.cpp): y -> bit 3i, x -> 3i+1, z -> 3i+2.
__host__ __device__ inline uint64_t sp_spread(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffULL;
    x = (x | x << 16) & 0x1f0000ff0000ffULL;
    x = (x | x << 8) & 0x100f00f00f00f00fULL;
    x = (x | x << 4) & 0x10c30c30c30c30c3ULL;
    x = (x | x << 2) & 0x1249249249249249ULL;
    return x;
}
__host__ __device__ inline uint32_t sp_compact(uint64_t x) {
    x &= 0x1249249249249249ULL;
    x = (x ^ (x >> 2)) & 0x10c30c30c30c30c3ULL;
    x = (x ^ (x >> 4)) & 0x100f00f00f00f00fULL;
    x = (x ^ (x >> 8)) & 0x1f0000ff0000ffULL;
    x = (x ^ (x >> 16)) & 0x1f00000000ffffULL;
    x = (x ^ (x >> 32)) & 0x1fffffULL;
    return (uint32_t)x;
}
__host__ __device__ inline uint64_t sp_key(int x, int y, int z) {
    return sp_spread((uint32_t)y) | (sp_spread((uint32_t)x) << 1) | (sp_spread((uint32_t)z) << 2);
}
__host__ __device__ inline void sp_unkey(uint64_t k, int& x, int& y, int& z) {
    y = (int)sp_compact(k);
    x = (int)sp_compact(k >> 1);
    z = (int)sp_compact(k >> 2);
}

// Base-27 cell index of the ternary tree (morton3, multipole_method_functions.
// cpp:49-86): level digit 9 z + 3 x + y, finest level least significant.
__host__ __device__ inline uint64_t sp_tkey(int x, int y, int z, int levels) {
    uint64_t k = 0, w = 1;
    for (int i = 0; i < levels; i++) {
        k += (uint64_t)(9 * (z % 3) + 3 * (x % 3) + (y % 3)) * w;
        x /= 3;
        y /= 3;
        z /= 3;
        w *= 27;
    }
    return k;
}
__host__ __device__ inline void sp_tunkey(uint64_t k, int& x, int& y, int& z) {
    x = y = z = 0;
    int w = 1;
    while (k) {
        const int d = (int)(k % 27);
        k /= 27;
        y += (d % 3) * w;
        x += ((d / 3) % 3) * w;
        z += (d / 9) * w;
        w *= 3;
    }
}
__host__ __device__ inline int sp_pow3(int l) {
    int w = 1;
    for (int i = 0; i < l; i++) w *= 3;
    return w;
}
// v in [-w, 2w) wrapped into [0, w); *shift = the cell image it came from
__device__ inline int sp_wrap(int v, int w, int& shift) {
    shift = v < 0 ? -1 : (v >= w ? 1 : 0);
    return v - shift * w;
}

// index of key among the nb sorted keys, -1 if absent
__device__ inline int sp_find(const uint64_t* __restrict__ keys, int nb, uint64_t key) {
    int lo = 0, hi = nb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < nb && keys[lo] == key) ? lo : -1;
}

// Geometry of the reference's grid: boxMin, rootBoxSize, the leaf box size
// (float as morton() computes it, double as p2m / l2p do), 2^maxLevel and the
// cells per axis; GL is the leaf level of the enclosing cube.
struct SpGeom {
    float bmin[3];
    float root;
    float bsf;
    double bsd;
    int side;
    int bdv[3];
    int GL;
    int rot;   // rotational symmetry: the quadrant map below the domain level
    int tern;  // translational: ternary tree over one periodic cell, base-27 keys
};

__device__ inline int sp_f2i(float v) { return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT_MIN; }

// Domain of a source (setDomainSize :714-737); false when outside the grid.
__device__ inline bool sp_domain(const SpGeom& g, const float4& p, int nd[3]) {
    const float c[3] = {p.x, p.y, p.z};
    bool in = true;
    for (int a = 0; a < 3; a++) {
        const float t = (c[a] - g.bmin[a]) / g.root;
        int v = sp_f2i(t);
        if (t < 0) v--;
        if (g.root == 0) v = 0;
        nd[a] = v;
        if (v < 0 || v >= g.bdv[a]) in = false;
    }
    return in;
}

// Centre of a leaf box as p2m / l2p form it: float corner of its domain plus
// (local index + 0.5) leaf sizes in double, stored as float.
__device__ inline void sp_leaf_centre(const SpGeom& g, uint64_t key, float cen[3]) {
    int gc[3];
    if (g.tern) sp_tunkey(key, gc[0], gc[1], gc[2]);
    else sp_unkey(key, gc[0], gc[1], gc[2]);
    for (int a = 0; a < 3; a++) {
        const int nd = gc[a] / g.side, nl = gc[a] % g.side;
        const float dmin = sp_madd_unfused(g.bmin[a], (float)nd, g.root);
        cen[a] = (float)sp_madd_unfused((double)dmin, nl + 0.5, g.bsd);
    }
}

__global__ void k_sp_flags(const int* __restrict__ st, int i0, int m, char* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        const int s = st[i0 + i];
        f[i] = (s == 1 || s == 2) ? 1 : 0;
    }
}

// Float bounds of the sources: per-block partials, then one block.
__global__ void k_sp_bbox(const float4* __restrict__ P, const int* __restrict__ src, const int* __restrict__ nsrc,
                          float* __restrict__ part) {
    __shared__ float sh[6][256];
    const int ns = *nsrc;
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        const float4 p = P[src[s]];
        mn[0] = fminf(mn[0], p.x); mx[0] = fmaxf(mx[0], p.x);
        mn[1] = fminf(mn[1], p.y); mx[1] = fmaxf(mx[1], p.y);
        mn[2] = fminf(mn[2], p.z); mx[2] = fmaxf(mx[2], p.z);
    }
    for (int a = 0; a < 3; a++) {
        sh[a][threadIdx.x] = mn[a];
        sh[3 + a][threadIdx.x] = mx[a];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int a = 0; a < 3; a++) {
                sh[a][threadIdx.x] = fminf(sh[a][threadIdx.x], sh[a][threadIdx.x + w]);
                sh[3 + a][threadIdx.x] = fmaxf(sh[3 + a][threadIdx.x], sh[3 + a][threadIdx.x + w]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ void k_sp_bbox_final(const float* __restrict__ part, int nparts, float* __restrict__ out6) {
    const int a = threadIdx.x;
    if (a >= 6) return;
    float v = a < 3 ? INFINITY : -INFINITY;
    for (int b = 0; b < nparts; b++) v = a < 3 ? fminf(v, part[b * 6 + a]) : fmaxf(v, part[b * 6 + a]);
    out6[a] = v;
}

// Domain histogram (setOptimumLevel counts per domain) and sources outside the grid.
constexpr int kSpHistLds = 4096;  // domain grids up to this size are counted in LDS first

// Domain histogram (setOptimumLevel counts per domain) and sources outside the
// grid (ctr[0]); a block counts into LDS when the grid is small, so a grid of a
// few domains does not serialise a million global atomics.
__global__ __launch_bounds__(256) void k_sp_domains(const float4* __restrict__ P, const int* __restrict__ src, int ns,
                                                    SpGeom g, int ndom, int* __restrict__ hist, int* __restrict__ ctr) {
    __shared__ int lh[kSpHistLds];
    const bool local = ndom <= kSpHistLds;
    if (local)
        for (int d = threadIdx.x; d < ndom; d += blockDim.x) lh[d] = 0;
    __syncthreads();
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < ns; s += gridDim.x * blockDim.x) {
        int nd[3];
        if (!sp_domain(g, P[src[s]], nd)) {
            atomicAdd(ctr, 1);
            continue;
        }
        const int d = (nd[0] * g.bdv[1] + nd[1]) * g.bdv[2] + nd[2];
        if (local) atomicAdd(&lh[d], 1);
        else atomicAdd(&hist[d], 1);
    }
    __syncthreads();
    if (local)
        for (int d = threadIdx.x; d < ndom; d += blockDim.x)
            if (lh[d]) atomicAdd(&hist[d], lh[d]);
}

// ctr[1] = the largest domain count.
__global__ void k_sp_hist_max(const int* __restrict__ hist, int ndom, int* __restrict__ ctr) {
    int m = 0;
    for (int d = blockIdx.x * blockDim.x + threadIdx.x; d < ndom; d += gridDim.x * blockDim.x) m = max(m, hist[d]);
    if (m) atomicMax(ctr + 1, m);
}

// Leaf key of each source (morton :1014-1040 in global cell coordinates).
__global__ void k_sp_keys(const float4* __restrict__ P, const int* __restrict__ src, int ns, SpGeom g,
                          uint64_t* __restrict__ keys, int* __restrict__ vals) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int i = src[s];
    const float4 p = P[i];
    int nd[3];
    sp_domain(g, p, nd);
    const float c[3] = {p.x, p.y, p.z};
    int gc[3];
    for (int a = 0; a < 3; a++) {
        const float dmin = sp_madd_unfused(g.bmin[a], (float)nd[a], g.root);
        int nl = sp_f2i((c[a] - dmin) / g.bsf);
        if (nl >= g.side) nl--;
        gc[a] = nd[a] * g.side + nl;
    }
    keys[s] = g.tern ? sp_tkey(gc[0], gc[1], gc[2], g.GL) : sp_key(gc[0], gc[1], gc[2]);
    vals[s] = i;
}

__global__ void k_sp_gather(const float4* __restrict__ P, const int* __restrict__ vals, int ns,
                            const uint64_t* __restrict__ keys, float4* __restrict__ sp, int* __restrict__ flag) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ns) return;
    sp[a] = P[vals[a]];
    flag[a] = (a > 0 && keys[a] != keys[a - 1]) ? 1 : 0;
}

__global__ void k_sp_shift(const uint64_t* __restrict__ in, int n, int tern, uint64_t* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < n) out[b] = tern ? in[b] / 27 : in[b] >> 3;
}

// ------------------------------------------------------------- operators --

// One dense operator per block, one real input component per thread (column).
// ops[o] = (kind, a, b, c) as sp_apply_op; swz != 0 stores the MFMA A-fragment
// order: element (row, col) at ((col/4) RT + row/16) 64 + (col%4) 16 + row%16.
__global__ __launch_bounds__(128) void k_sp_build_ops(SphTab t, int p, int KP, const int4* __restrict__ ops,
                                                      double h, int swz, double* __restrict__ T) {
    const int o = blockIdx.x, col = threadIdx.x;
    if (col >= KP) return;
    const int4 d = ops[o];
    double* To = T + (size_t)o * KP * KP;
    const int nc2 = p * (p + 1);
    const int RT = KP / 16;
    double2 X[kSpNcMax], Y[kSpNcMax];
    if (col < nc2 && d.x >= 0) sp_apply_op(t, p, d.x, d.y, d.z, d.w, h, col, X, Y);
    for (int row = 0; row < KP; row++) {
        double v = 0.0;
        if (col < nc2 && row < nc2 && d.x >= 0) v = (row & 1) ? Y[row >> 1].y : Y[row >> 1].x;
        const size_t at = swz ? ((size_t)((col >> 2) * RT + (row >> 4)) * 64 + (col & 3) * 16 + (row & 15))
                              : (size_t)row * KP + col;
        To[at] = v;
    }
}

// P2M (p2m_kernel, fmm_cpukernel.cpp:281-384): one block of 64 lanes per leaf
// box.  Lane l serves order m = l % p of the particles l / p, l / p + G, ...
// (G = 64 / p particle groups), running the reference's Legendre recurrence
// of that order with its running powers of rho; the G group sums of each
// coefficient are then added in group order.
__global__ __launch_bounds__(64) void k_sp_p2m(SphTab t, int p, int KP, SpGeom g, const uint64_t* __restrict__ keys,
                                               const int* __restrict__ first, const int* __restrict__ count,
                                               const float4* __restrict__ sp, double* __restrict__ M) {
    __shared__ double2 part[64][kSpPmax];
    const int b = blockIdx.x, lane = threadIdx.x;
    const int G = 64 / p, m = lane % p, grp = lane / p;
    double* Mb = M + (size_t)b * KP;
    double2 acc[kSpPmax];
    for (int n = 0; n < kSpPmax; n++) acc[n] = make_double2(0, 0);
    if (grp < G) {
        float cen[3];
        sp_leaf_centre(g, keys[b], cen);
        for (int a = first[b] + grp; a < first[b] + count[b]; a += G) {
            const float4 q = sp[a];
            const float fx = q.x - cen[0], fy = q.y - cen[1], fz = q.z - cen[2];
            double rho, alpha, beta;
            sp_cart2sph(rho, alpha, beta, fx, fy, fz);
            rho /= g.bsd;
            const double xx = cos(alpha);
            const double s2 = sqrt((1 - xx) * (1 + xx));
            double fact = 1, pn = 1, rhom = 1.0;
            for (int mm = 0; mm < m; mm++) {  // the reference's running P_m^m and rho^m
                rhom *= rho;
                pn = -pn * fact * s2;
                fact += 2;
            }
            double pv = pn;
            const double eb = -m * beta;
            const double ce = cos(eb), se = sin(eb);
            double Y = rhom * (double)t.factorial[m * m + 2 * m] * pv;
            acc[m].x += (double)q.w * Y * ce;
            acc[m].y += (double)q.w * Y * se;
            double p1 = pv;
            pv = xx * (2 * m + 1) * pv;
            double rhon = rhom * rho;
            for (int n = m + 1; n < p; n++) {
                Y = rhon * (double)t.factorial[n * n + n + m] * pv;
                acc[n].x += (double)q.w * Y * ce;
                acc[n].y += (double)q.w * Y * se;
                const double p2 = p1;
                p1 = pv;
                pv = (xx * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
                rhon *= rho;
            }
        }
    }
    for (int n = 0; n < kSpPmax; n++) part[lane][n] = acc[n];
    __syncthreads();
    // lane c < p(p+1)/2 adds coefficient (n, mc) over the groups in order
    const int nc = p * (p + 1) / 2;
    for (int c = lane; c < nc; c += 64) {
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= c) n++;
        const int mc = c - n * (n + 1) / 2;
        double2 v = make_double2(0, 0);
        for (int gg = 0; gg < G; gg++) {
            v.x += part[gg * p + mc][n].x;
            v.y += part[gg * p + mc][n].y;
        }
        Mb[2 * c] = v.x;
        Mb[2 * c + 1] = v.y;
    }
    for (int c = 2 * nc + lane; c < KP; c += 64) Mb[c] = 0.0;
}

// M2M: each parent sums its children (key order) through the child-octant
// operators; zop != 0: through the rotational operator of the child's z bit;
// tern: the 27 ternary children (parent key = key / 27, operator key % 27).
__global__ __launch_bounds__(128) void k_sp_m2m(const double* __restrict__ Top, int KP, int nc2, int zop, int tern,
                                                const uint64_t* __restrict__ ckeys, int ncb, const double* __restrict__ Mc,
                                                const uint64_t* __restrict__ pkeys, double* __restrict__ Mp) {
    const int pb = blockIdx.x, row = threadIdx.x;
    if (row >= KP) return;
    const uint64_t pk = pkeys[pb];
    int lo = 0, hi = ncb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((tern ? ckeys[mid] / 27 : ckeys[mid] >> 3) < pk) lo = mid + 1;
        else hi = mid;
    }
    double acc = 0.0;
    if (row < nc2)
        for (int c = lo; c < ncb && (tern ? ckeys[c] / 27 : ckeys[c] >> 3) == pk; c++) {
            const unsigned slot = tern  ? (unsigned)(ckeys[c] % 27)
                                  : zop ? (unsigned)((ckeys[c] >> 2) & 1u)
                                        : (unsigned)(ckeys[c] & 7u);
            const double* T = Top + ((size_t)slot * KP + row) * KP;
            const double* m = Mc + (size_t)c * KP;
            double s = 0.0;
            for (int k = 0; k < nc2; k++) s += T[k] * m[k];
            acc += s;
        }
    Mp[(size_t)pb * KP + row] = acc;
}

// L2L: each child takes its parent's local expansion through its octant's
// (tern: ternary child's) operator.
__global__ __launch_bounds__(128) void k_sp_l2l(const double* __restrict__ Top, int KP, int nc2, int tern,
                                                const uint64_t* __restrict__ pkeys, int npb, const double* __restrict__ Lp,
                                                const uint64_t* __restrict__ ckeys, double* __restrict__ Lc) {
    const int cb = blockIdx.x, row = threadIdx.x;
    if (row >= KP) return;
    const uint64_t ck = ckeys[cb];
    const int pb = sp_find(pkeys, npb, tern ? ck / 27 : ck >> 3);
    double s = 0.0;
    if (row < nc2 && pb >= 0) {
        const double* T = Top + ((size_t)(tern ? ck % 27 : ck & 7u) * KP + row) * KP;
        const double* l = Lp + (size_t)pb * KP;
        for (int k = 0; k < nc2; k++) s += T[k] * l[k];
    }
    Lc[(size_t)cb * KP + row] = s;
}

typedef double sp_d4 __attribute__((ext_vector_type(4)));

// M2L of one level on the FP64 matrix cores.  Block = one wave = 16 target
// boxes of one child octant (tiles[b] = octant, first, count); lane l serves
// target l % 16 and k-slice l / 16 of every 16x16x4 product.  For each of the
// octant's 189 offsets (offs[octant][o] = source - target, operator slot) the
// wave gathers the 16 source expansions (zero where the source box is empty)
// and accumulates T(slot) x M.  Fragments of v_mfma_f64_16x16x4f64 (measured
// on gfx950): lane l holds A[l % 16][l / 16], B[l / 16][l % 16], and D rows
// l / 16 + 4 r (r = 0..3) of column l % 16.  ctr[0] counts the box pairs that exist.
// jd > 0 (rotational, below the domain level): a source cell of the window of
// jd cells per axis in x, y is mapped to the real cell it images and its
// multipole is gathered times i^(-m k) (m2l_kernel_rotations,
// fmm_cpukernel_rotational.cpp:367-372); mord[c] = order m of real component c.
// cells (tern, may be null): box of each cell of the level, (x W + y) W + z,
// -1 where empty; replaces the binary search of the level's keys.
// tern: the ternary tree of one periodic cell; offs holds the 702 offsets of
// each of the 27 child positions and a source cell outside the 3^level cells
// per axis is its periodic image (getInteractionListTranslations M2L_Subcell,
// multipole_method_force_calculation_translational.cpp:1725-1834).
// Work unit = (tile, chunk of the offset list): with nchunk > 1 each unit
// writes its 16 x KP partial sums to part[chunk][tile] and k_sp_m2l_reduce adds
// them to L in chunk order, so a level with few tiles still fills the chip.
template <int RT>
__global__ __launch_bounds__(64) void k_sp_m2l(const double* __restrict__ T, const double* __restrict__ M,
                                               double* __restrict__ L, const uint64_t* __restrict__ keys, int nb,
                                               int level, int jd, int tern, int noff, const int* __restrict__ grp,
                                               const int4* __restrict__ tiles, int ntiles, int nchunk,
                                               double* __restrict__ part, const int4* __restrict__ offs,
                                               const signed char* __restrict__ mord, const int* __restrict__ cells,
                                               unsigned long long* __restrict__ ctr) {
    constexpr int KP = RT * 16, KS = KP / 4;
    const int lane = threadIdx.x, j = lane & 15, kq = lane >> 4;
    const int tile = blockIdx.x / nchunk, chunk = blockIdx.x % nchunk;
    const int o0 = chunk * noff / nchunk, o1 = (chunk + 1) * noff / nchunk;
    const int4 tl = tiles[tile];
    const int tb = j < tl.z ? grp[tl.y + j] : -1;
    int tx = 0, ty = 0, tz = 0;
    if (tb >= 0) {
        if (tern) sp_tunkey(keys[tb], tx, ty, tz);
        else sp_unkey(keys[tb], tx, ty, tz);
    }
    const int W = tern ? sp_pow3(level) : 1 << level;
    sp_d4 acc[RT];
    for (int r = 0; r < RT; r++) acc[r] = (sp_d4){0.0, 0.0, 0.0, 0.0};
    const int4* od = offs + tl.x * noff;
    unsigned long long pairs = 0;
    for (int o = o0; o < o1; o++) {
        const int4 d = od[o];
        int sb = -1, quad = 0;
        if (tb >= 0 && tern) {
            int e0, e1, e2;
            const int sx = sp_wrap(tx + d.x, W, e0), sy = sp_wrap(ty + d.y, W, e1), sz = sp_wrap(tz + d.z, W, e2);
            sb = cells ? cells[((size_t)sx * W + sy) * W + sz] : sp_find(keys, nb, sp_tkey(sx, sy, sz, level));
        } else if (tb >= 0) {
            int sx = tx + d.x, sy = ty + d.y;
            const int sz = tz + d.z;
            int lim = W;
            if (jd > 0) {
                lim = jd;
                quad = sp_quadrant(tx + d.x, ty + d.y, jd, sx, sy);
            }
            if (sx >= 0 && sy >= 0 && sz >= 0 && sx < lim && sy < lim && sz < W)
                sb = sp_find(keys, nb, sp_key(sx, sy, sz));
        }
        const unsigned long long have = __ballot(sb >= 0);
        if (have == 0) continue;
        pairs += __popcll(have & 0xffffull);
        const double* Ms = M + (size_t)(sb >= 0 ? sb : 0) * KP;
        const double* To = T + (size_t)d.w * KP * KP;
        for (int s = 0; s < KS; s++) {
            double bv = 0.0;
            if (sb >= 0) {
                const int c = 4 * s + kq;
                if (quad == 0) {
                    bv = Ms[c];
                } else {
                    const double2 z = sp_phase(make_double2(Ms[c & ~1], Ms[c | 1]), mord[c], quad);
                    bv = (c & 1) ? z.y : z.x;
                }
            }
            for (int r = 0; r < RT; r++) {
                const double av = To[(size_t)(s * RT + r) * 64 + lane];
                acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[r], 0, 0, 0);
            }
        }
    }
    if (lane == 0 && pairs) atomicAdd(ctr, pairs);
    if (nchunk == 1) {
        if (tb >= 0) {
            double* Lt = L + (size_t)tb * KP;
            for (int r = 0; r < RT; r++)
                for (int i = 0; i < 4; i++) Lt[16 * r + kq + 4 * i] += acc[r][i];  // D row kq + 4 i, column j
        }
    } else {
        double* Pt = part + ((size_t)chunk * ntiles + tile) * 16 * KP + (size_t)j * KP;
        for (int r = 0; r < RT; r++)
            for (int i = 0; i < 4; i++) Pt[16 * r + kq + 4 * i] = acc[r][i];
    }
}

// L[target] += sum over chunks (in order) of the M2L partials; one thread per
// (tile, target column, row).
__global__ void k_sp_m2l_reduce(const double* __restrict__ part, int ntiles, int nchunk, int KP,
                                const int4* __restrict__ tiles, const int* __restrict__ grp, double* __restrict__ L) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)ntiles * 16 * KP) return;
    const int tile = (int)(i / (16 * KP)), j = (int)((i / KP) % 16), row = (int)(i % KP);
    const int4 tl = tiles[tile];
    if (j >= tl.z) return;
    double v = 0.0;
    for (int c = 0; c < nchunk; c++) v += part[((size_t)c * ntiles + tile) * 16 * KP + (size_t)j * KP + row];
    L[(size_t)grp[tl.y + j] * KP + row] += v;
}

// The 27 neighbour leaves of leaf b in the reference's order (x, then y, then
// z): 4 * leaf + quadrant (the quadrant map of the rotational lists, leaf
// window of side cells per axis; 0 without symmetry), -1 where absent.
// tern: 27 * leaf + image code (sx + 1) 9 + (sy + 1) 3 + (sz + 1) of the
// periodic wrap (getInteractionListTranslations P2P, multipole_method_force_
// calculation_translational.cpp:1217-1334).
__device__ inline void sp_neighbours(const SpGeom& g, const uint64_t* __restrict__ keys, int nb, int b, int* nbr) {
    int ix, iy, iz;
    if (g.tern) {
        sp_tunkey(keys[b], ix, iy, iz);
        for (int k = 0; k < 27; k++) {
            int e0, e1, e2;
            const int jx = sp_wrap(ix + k / 9 - 1, g.side, e0), jy = sp_wrap(iy + (k / 3) % 3 - 1, g.side, e1),
                      jz = sp_wrap(iz + k % 3 - 1, g.side, e2);
            const int f = sp_find(keys, nb, sp_tkey(jx, jy, jz, g.GL));
            nbr[k] = f >= 0 ? 27 * f + (e0 + 1) * 9 + (e1 + 1) * 3 + (e2 + 1) : -1;
        }
        return;
    }
    sp_unkey(keys[b], ix, iy, iz);
    const int W = 1 << g.GL;
    for (int k = 0; k < 27; k++) {
        int jx = ix + k / 9 - 1, jy = iy + (k / 3) % 3 - 1, quad = 0, lim = W;
        const int jz = iz + k % 3 - 1;
        if (g.rot) {
            lim = g.side;
            quad = sp_quadrant(ix + k / 9 - 1, iy + (k / 3) % 3 - 1, lim, jx, jy);
        }
        const int f = (jx >= 0 && jy >= 0 && jz >= 0 && jx < lim && jy < lim && jz < W)
                          ? sp_find(keys, nb, sp_key(jx, jy, jz))
                          : -1;
        nbr[k] = f >= 0 ? 4 * f + quad : -1;
    }
}

// P2P work units: every leaf's targets in runs of 64.  ucnt[b] = runs of leaf
// b; ucost[b] = 64 x its near-field sources (the sort key of its units);
// ctr[1] counts the box pairs, ctr[2] the ion pairs.
__global__ void k_sp_unit_counts(SpGeom g, const uint64_t* __restrict__ keys, int nb, const int* __restrict__ count,
                                 int* __restrict__ ucnt, unsigned* __restrict__ ucost,
                                 unsigned long long* __restrict__ ctr) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int nbr[27];
    sp_neighbours(g, keys, nb, b, nbr);
    long near = 0;
    int nnb = 0;
    for (int k = 0; k < 27; k++)
        if (nbr[k] >= 0) {
            near += count[g.tern ? nbr[k] / 27 : nbr[k] >> 2];
            nnb++;
        }
    atomicAdd(ctr + 1, (unsigned long long)nnb);
    atomicAdd(ctr + 2, (unsigned long long)count[b] * (unsigned long long)near);
    ucnt[b] = (count[b] + 63) / 64;
    ucost[b] = (unsigned)min(64L * near, 0xffffffffL);
}

// Unit u of leaf b at uoff[b] + u: value b << 32 | u, key ~cost so that an
// ascending sort puts the longest units first.
__global__ void k_sp_unit_fill(int nb, const int* __restrict__ ucnt, const int* __restrict__ uoff,
                               const unsigned* __restrict__ ucost, unsigned* __restrict__ ukey,
                               unsigned long long* __restrict__ uval) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    for (int u = 0; u < ucnt[b]; u++) {
        ukey[uoff[b] + u] = ~ucost[b];
        uval[uoff[b] + u] = ((unsigned long long)b << 32) | (unsigned)u;
    }
}

// P2P (p2p_kernel, fmm_cpukernel.cpp:194-219, over DirectForceFunction::direct,
// direct_force_method.cpp:153-190): one wave per unit of 64 targets of one
// leaf, longest units first.  For each of the 27 neighbour leaves in the
// reference's order the wave stages the leaf's sources in LDS, 256 at a time
// (coalesced loads), every lane sums its target's pairs over the leaf in
// FP64 (float displacement, as the reference forms it) and adds K q / m times
// the leaf's sum.  A leaf imaged into quadrant k has its sources rotated by
// k pi / 2 as direct_rotational rotates them (direct_force_method_rotational.
// cpp:59-75: the float products by cos and sin in {0, 1, -1} are exact); a
// periodic image leaf shifts the target by -shift * cell in float as
// direct_translational forms it (direct_force_method_translational.cpp:324-345).
// near[c * ns + a] for the source at key-order position a (status 1 only).
__global__ __launch_bounds__(64) void k_sp_p2p(SpGeom g, double kc, const uint64_t* __restrict__ keys, int nb,
                                               const int* __restrict__ first, const int* __restrict__ count,
                                               const unsigned long long* __restrict__ units,
                                               const float4* __restrict__ sp,
                                               const int* __restrict__ sorig, int ns, const int* __restrict__ st,
                                               const int* __restrict__ spc, double* __restrict__ near) {
    __shared__ float4 tile[kSpP2pTile];
    __shared__ int nbr[27];
    const int lane = threadIdx.x;
    const unsigned long long uv = units[blockIdx.x];
    const int b = (int)(uv >> 32), u = (int)(uv & 0xffffffffu);
    if (lane == 0) sp_neighbours(g, keys, nb, b, nbr);
    __syncthreads();
    const int a = first[b] + 64 * u + lane;
    const bool tgt = a < first[b] + count[b] && st[sorig[a]] == 1;
    float4 pi = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    double ac = 0.0;
    if (tgt) {
        pi = sp[a];
        ac = kc * (double)pi.w / species_mass(spc[sorig[a]]);
    }
    double tot[3] = {0, 0, 0};
    for (int k = 0; k < 27; k++) {
        if (nbr[k] < 0) continue;
        const int sb = g.tern ? nbr[k] / 27 : nbr[k] >> 2, quad = g.tern ? 0 : nbr[k] & 3;
        const int code = g.tern ? nbr[k] % 27 : 13;
        const float ox = (float)(code / 9 - 1) * g.root, oy = (float)((code / 3) % 3 - 1) * g.root,
                    oz = (float)(code % 3 - 1) * g.root;
        const float px = pi.x - ox, py = pi.y - oy, pz = pi.z - oz;
        double ai[3] = {0, 0, 0};
        const int e0 = first[sb], e1 = first[sb] + count[sb];
        for (int s0 = e0; s0 < e1; s0 += kSpP2pTile) {
            __syncthreads();
            for (int t = lane; t < kSpP2pTile; t += 64) {
                float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (s0 + t < e1) {
                    q = sp[s0 + t];
                    const float x = q.x, y = q.y;
                    if (quad == 1) {
                        q.x = -y;
                        q.y = x;
                    } else if (quad == 2) {
                        q.x = -x;
                        q.y = -y;
                    } else if (quad == 3) {
                        q.x = y;
                        q.y = -x;
                    }
                }
                tile[t] = q;
            }
            __syncthreads();
            // the self pair (zero displacement) adds exactly zero, as the
            // reference's skip of j == i does; entries past the leaf have zero charge
            const int n = min(kSpP2pTile, e1 - s0);
            if (tgt) {
#pragma unroll 4
                for (int e = 0; e < n; e++) {
                    const float4 pj = tile[e];
                    const double d0 = (double)(px - pj.x), d1 = (double)(py - pj.y), d2 = (double)(pz - pj.z);
                    const double invD = rsqrt_f64(d0 * d0 + d1 * d1 + d2 * d2 + kSoftening);
                    const double invD3 = (double)pj.w * invD * invD * invD;
                    ai[0] += d0 * invD3;
                    ai[1] += d1 * invD3;
                    ai[2] += d2 * invD3;
                }
            }
        }
        for (int c = 0; c < 3; c++) tot[c] += ac * ai[c];
    }
    if (tgt)
        for (int c = 0; c < 3; c++) near[(size_t)c * ns + a] = tot[c];
}

// L2P (l2p_kernel, fmm_cpukernel.cpp:1188-1317) per target in key order, added
// to its P2P sum: out[c][i - i0] = P2P - K q / m * L2P.
__global__ __launch_bounds__(64) void k_sp_leaf(SphTab t, int p, int KP, SpGeom g, double kc,
                                                const uint64_t* __restrict__ keys, const int* __restrict__ sbox,
                                                const float4* __restrict__ sp, const int* __restrict__ sorig, int ns,
                                                const int* __restrict__ st, const int* __restrict__ spc,
                                                const double* __restrict__ near, const double* __restrict__ L, int i0,
                                                int m, double* __restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= ns) return;
    const int i = sorig[a];
    const int b = sbox[a];
    if (st[i] != 1) return;
    const float4 pi = sp[a];
    const double mass = species_mass(spc[i]);
    const double ac = kc * (double)pi.w / mass;
    const double tot[3] = {near[a], near[(size_t)ns + a], near[2 * (size_t)ns + a]};
    // L2P
    float cen[3];
    sp_leaf_centre(g, keys[b], cen);
    const float fx = pi.x - cen[0], fy = pi.y - cen[1], fz = pi.z - cen[2];
    double r, theta, phi;
    sp_cart2sph(r, theta, phi, fx, fy, fz);
    const double xx = cos(theta), yy = sin(theta);
    const double s2 = sqrt((1 - xx) * (1 + xx));
    double Y[kSpNe2], Yt[kSpNe2];
    double fact = 1, pn = 1;
    for (int mm = 0; mm < p; mm++) {
        double pv = pn;
        int nm = mm * mm + 2 * mm;
        Y[nm] = (double)t.factorial[nm] * pv;
        double p1 = pv;
        pv = xx * (2 * mm + 1) * pv;
        Yt[nm] = (double)t.factorial[nm] * (pv - (mm + 1) * xx * p1) / yy;
        for (int n = mm + 1; n < p; n++) {
            nm = n * n + n + mm;
            Y[nm] = (double)t.factorial[nm] * pv;
            const double p2 = p1;
            p1 = pv;
            pv = (xx * (2 * n + 1) * p1 - (n + mm) * p2) / (n - mm + 1);
            Yt[nm] = (double)t.factorial[nm] * ((n - mm + 1) * pv - (n + 1) * xx * p1) / yy;
        }
        pn = -pn * fact * s2;
        fact += 2;
    }
    const double* Lb = L + (size_t)b * KP;
    double aR = 0, aT = 0, aP = 0, rn = 1;
    r /= g.bsd;
    for (int n = 0; n < p; n++) {
        int nm = n * n + n, nms = n * (n + 1) / 2;
        const double rr = n * rn / r * Y[nm];
        const double rt = rn * Yt[nm];
        aR += rr * Lb[2 * nms];
        aT += rt * Lb[2 * nms];
        for (int mm = 1; mm <= n; mm++) {
            nm = n * n + n + mm;
            nms = n * (n + 1) / 2 + mm;
            const double ce = cos(mm * phi), se = sin(mm * phi);
            const double vr = n * rn / r * Y[nm], vt = rn * Yt[nm], vp = mm * rn * Y[nm];
            const double lr = Lb[2 * nms], li = Lb[2 * nms + 1];
            aR += 2 * (ce * vr * lr - se * vr * li);
            aT += 2 * (ce * vt * lr - se * vt * li);
            aP += 2 * ((-(se * vp)) * lr - (ce * vp) * li);
        }
        rn *= r;
    }
    const double h2 = g.bsd * g.bsd;
    aR /= h2;
    aT /= h2;
    aP /= h2;
    const double st_ = sin(theta), ct = cos(theta), sf = sin(phi), cf = cos(phi);
    const double f[3] = {st_ * cf * aR + ct * cf / r * aT - sf / r / st_ * aP,
                         st_ * sf * aR + ct * sf / r * aT + cf / r / st_ * aP, ct * aR - st_ / r * aT};
    const int row = i - i0;
    for (int c = 0; c < 3; c++) out[(size_t)c * m + row] = tot[c] - ac * f[c];
}

__global__ void k_sp_zero(double* __restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ void k_sp_zero_i(int* __restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_sp_fill_i(int* __restrict__ a, size_t n, int v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// cells[(x W + y) W + z] = b for every box b of a ternary level of W cells per axis.
__global__ void k_sp_cells_set(const uint64_t* __restrict__ keys, int nb, int W, int* __restrict__ cells) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int x, y, z;
    sp_tunkey(keys[b], x, y, z);
    if (x < W && y < W && z < W) cells[((size_t)x * W + y) * W + z] = b;
}

__global__ void k_sp_octants(const uint64_t* __restrict__ keys, int nb, int tern, unsigned* __restrict__ okey,
                             int* __restrict__ oval, int* __restrict__ ocnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const unsigned o = tern ? (unsigned)(keys[b] % 27) : (unsigned)(keys[b] & 7u);
    okey[b] = o;
    oval[b] = b;
    atomicAdd(&ocnt[o], 1);
}

__global__ void k_sp_from_aos(const float4* __restrict__ vel, const float4* __restrict__ acc, int n,
                              int* __restrict__ st, int* __restrict__ sp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        st[i] = (int)acc[i].w;
        sp[i] = (int)vel[i].w;
    }
}

__global__ void k_sp_interleave(const double* __restrict__ a, int m, double* __restrict__ out3) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        out3[3 * (size_t)i] = a[i];
        out3[3 * (size_t)i + 1] = a[m + i];
        out3[3 * (size_t)i + 2] = a[2 * (size_t)m + i];
    }
}

inline int sp_bdiv(long a, int b) { return (int)((a + b - 1) / b); }

// setOptimumLevel's level_switch (multipole_method_force_calculation.cpp:842-869).
int sp_level_of_count(int c) {
    static const float sw[8] = {200.0f, 1000.0f, 1.3e4f, 1e5f, 7e5f, 5e6f, 3e7f, 1.5e8f};
    for (int i = 0; i < 8; i++)
        if ((float)c < sw[i]) return i;
    return 8;
}

int sp_state(shp_handle h, SphState** out) {
    if (h->sph) {
        *out = static_cast<SphState*>(h->sph);
        return SHP_OK;
    }
    const SphHostTables* T = sp_host_tables();
    SphState* S = new SphState();
    const size_t bytes = sizeof(SphHostTables);
    hipError_t e = hipMalloc(&S->buf, bytes);
    if (e == hipSuccess) e = hipMemcpy(S->buf, T, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (S->buf) (void)hipFree(S->buf);
        delete S;
        set_error("spherical FMM tables: %s", hipGetErrorString(e));
        return SHP_ERR_HIP;
    }
    const SphHostTables* D = static_cast<const SphHostTables*>(S->buf);
    S->tab.factorial = D->factorial;
    S->tab.anm = D->anm;
    S->tab.Anm = D->Anm;
    S->tab.Ynm = D->Ynm;
    S->tab.Dnm = &D->Dnm[0][0][0];
    S->tab.link = D->link;
    h->sph = S;
    *out = S;
    return SHP_OK;
}

// The 189 offsets (source - target) of a target in child octant o, in the
// reference's list order (parent neighbour, then child; x, y, z), each with the
// operator slot of target - source.
void sp_offsets(std::vector<int4>& offs) {
    offs.clear();
    for (int o = 0; o < 8; o++) {
        const int c[3] = {(o >> 1) & 1, o & 1, (o >> 2) & 1};
        for (int ex = -1; ex <= 1; ex++)
            for (int ey = -1; ey <= 1; ey++)
                for (int ez = -1; ez <= 1; ez++)
                    for (int qx = 0; qx < 2; qx++)
                        for (int qy = 0; qy < 2; qy++)
                            for (int qz = 0; qz < 2; qz++) {
                                const int dx = 2 * ex + qx - c[0], dy = 2 * ey + qy - c[1], dz = 2 * ez + qz - c[2];
                                if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                                const int slot = (-dx + 3) * 49 + (-dy + 3) * 7 + (-dz + 3);
                                offs.push_back(make_int4(dx, dy, dz, slot));
                            }
    }
}

// The 702 offsets (source - target) of a ternary cell at child position c
// (base-27 digit, c = 9 z + 3 x + y), in the list order of M2L_Subcell
// (parent neighbour, then child; x, y, z), each with the operator slot of
// target - source in [-5, 5]^3.
void sp_offsets_tern(std::vector<int4>& offs) {
    offs.clear();
    for (int o = 0; o < 27; o++) {
        const int c[3] = {(o / 3) % 3, o % 3, o / 9};
        for (int ex = -1; ex <= 1; ex++)
            for (int ey = -1; ey <= 1; ey++)
                for (int ez = -1; ez <= 1; ez++)
                    for (int qx = 0; qx < 3; qx++)
                        for (int qy = 0; qy < 3; qy++)
                            for (int qz = 0; qz < 3; qz++) {
                                const int dx = 3 * ex + qx - c[0], dy = 3 * ey + qy - c[1], dz = 3 * ez + qz - c[2];
                                if (std::abs(dx) <= 1 && std::abs(dy) <= 1 && std::abs(dz) <= 1) continue;
                                const int slot = (-dx + 5) * 121 + (-dy + 5) * 11 + (-dz + 5);
                                offs.push_back(make_int4(dx, dy, dz, slot));
                            }
    }
}

template <int RT>
void sp_launch_m2l(hipStream_t s, int ntiles, int nchunk, double* part, const double* T, const double* M, double* L,
                   const uint64_t* keys, int nb, int level, int jd, int tern, int noff, const int* grp,
                   const int4* tiles, const int4* offs, const signed char* mord, const int* cells,
                   unsigned long long* ctr) {
    hipLaunchKernelGGL(k_sp_m2l<RT>, dim3(ntiles * nchunk), dim3(64), 0, s, T, M, L, keys, nb, level, jd, tern, noff,
                       grp, tiles, ntiles, nchunk, part, offs, mord, cells, ctr);
    if (nchunk > 1)
        hipLaunchKernelGGL(k_sp_m2l_reduce, dim3(sp_bdiv((long)ntiles * 16 * RT * 16, 256)), dim3(256), 0, s, part,
                           ntiles, nchunk, RT * 16, tiles, grp, L);
}

// Evaluates targets (status 1) in [i0, i1) from the sources (status 1 or 2) in
// the same range; out[c * m + i - i0] for m = i1 - i0.  rot: the rotational FMM;
// cell > 0: the translational FMM over the periodic cube of that side.
int sph_eval(shp_handle h, const float4* P, const int* st, const int* spc, int i0, int i1, int p, int forced,
             int rot, float cell, double* out) {
    hipStream_t stream = h->stream;
    const int m = i1 - i0;
    h->sph_have_stats = false;
    hipLaunchKernelGGL(k_sp_zero, dim3(sp_bdiv(3L * m, 256)), dim3(256), 0, stream, out, (size_t)3 * m);
    if (m == 0) return SHP_OK;
    SphState* S = nullptr;
    int rc = sp_state(h, &S);
    if (rc) return rc;
    const int nc2 = p * (p + 1);
    const int KP = (nc2 + 15) / 16 * 16;
    const int RT = KP / 16;
    // sources and their float bounds
    SHP_WS(flags, char, h, WS_FLAGS, m);
    SHP_WS(misc, int, h, WS_COUNTS, 128);
    int* nsel = misc;                                       // [0]
    int* dctr = misc + 2;                                   // [2] outside, [3] largest domain count
    int* nruns = misc + 4;                                  // [4]
    int* ocnt = misc + 64;                                  // [64..90]: boxes per child position
    float* box6 = (float*)(misc + 16);                      // [16..21]
    unsigned long long* pctr = (unsigned long long*)(misc + 24);  // [24..29]: M2L pairs, P2P box and ion pairs
    hipLaunchKernelGGL(k_sp_flags, dim3(sp_bdiv(m, 256)), dim3(256), 0, stream, st, i0, m, flags);
    hipLaunchKernelGGL(k_sp_zero_i, dim3(1), dim3(128), 0, stream, misc, (size_t)128);
    // WS_SPH_KEYS: keys in/out (2 m u64), vals in/out (2 m int), selected sources (m int)
    char* kb = static_cast<char*>(h->get(WS_SPH_KEYS, (size_t)m * (2 * sizeof(uint64_t) + 3 * sizeof(int))));
    if (!kb) return SHP_ERR_NOMEM;
    uint64_t* kin = (uint64_t*)kb;
    uint64_t* kout = kin + m;
    int* vin = (int*)(kout + m);
    int* vout = vin + m;
    int* sel = vout + m;
    hipcub::CountingInputIterator<int> it(i0);
    size_t tb = 0;
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(nullptr, tb, it, flags, sel, nsel, m, stream));
    SHP_WS(cub, char, h, WS_CUB, tb);
    SHP_HIP_CHECK(hipcub::DeviceSelect::Flagged(cub, tb, it, flags, sel, nsel, m, stream));
    const int nbp = std::min(sp_bdiv(m, 256), 512);
    SHP_WS(part, float, h, WS_BBOX, 6 * (size_t)nbp);
    hipLaunchKernelGGL(k_sp_bbox, dim3(nbp), dim3(256), 0, stream, P, sel, nsel, part);
    hipLaunchKernelGGL(k_sp_bbox_final, dim3(1), dim3(64), 0, stream, part, nbp, box6);
    int ns = 0;
    float bb[6];
    float4 seed4;
    SHP_HIP_CHECK(hipMemcpyAsync(&ns, nsel, sizeof(int), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(bb, box6, sizeof(bb), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(&seed4, P + i0, sizeof(seed4), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    if (ns == 0) return SHP_OK;
    // the reference's domain grid (computeDomainBounds, host)
    SpGeom g{};
    const int tern = cell > 0 ? 1 : 0;
    if (tern) {
        // computeDomainBoundsTranslations (domain_bounds_functions.cpp:637-661): one
        // domain, boxMin = -rootBoxSize / 2 on every periodic axis
        for (int a = 0; a < 3; a++) {
            g.bmin[a] = -cell / 2;
            g.bdv[a] = 1;
        }
        g.root = cell;
        g.tern = 1;
    } else {
        const float seed[3] = {seed4.x, seed4.y, seed4.z};
        float dims[3], minO[3], maxO[3];
        domain_bounds_host(bb, bb + 3, seed, ns, 0, rot, nullptr, g.bmin, dims, g.bdv, minO, maxO);
        g.root = dims[0];
        g.rot = rot;
    }
    if (rot && (g.bdv[0] != 1 || g.bdv[1] != 1)) {
        set_error("shp_fmm_sph: rotational grid of %d x %d x %d domains (the reference's quadrant map works in "
                  "domain-local x, y: one domain per x and y)", g.bdv[0], g.bdv[1], g.bdv[2]);
        return SHP_ERR_UNSUPPORTED;
    }
    const long ndom = (long)g.bdv[0] * g.bdv[1] * g.bdv[2];
    if (ndom < 1 || ndom > kSpMaxDomains) {
        set_error("shp_fmm_sph: %ld domains (limit %d)", ndom, kSpMaxDomains);
        return SHP_ERR_UNSUPPORTED;
    }
    SHP_WS(hist, int, h, WS_SPH_LVL, ndom);
    hipLaunchKernelGGL(k_sp_zero_i, dim3(sp_bdiv(ndom, 256)), dim3(256), 0, stream, hist, (size_t)ndom);
    hipLaunchKernelGGL(k_sp_domains, dim3(std::min(sp_bdiv(ns, 256), 4 * h->num_cus)), dim3(256), 0, stream, P, sel,
                       ns, g, (int)ndom, hist, dctr);
    hipLaunchKernelGGL(k_sp_hist_max, dim3(std::min(sp_bdiv(ndom, 256), 256)), dim3(256), 0, stream, hist, (int)ndom,
                       dctr);
    int dres[2];
    SHP_HIP_CHECK(hipMemcpyAsync(dres, dctr, sizeof(dres), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    if (dres[0] > 0) {
        set_error("shp_fmm_sph: %d sources outside the domain grid (the reference's outlier passes)", dres[0]);
        return SHP_ERR_UNSUPPORTED;
    }
    const int maxLevel = forced != -2 ? forced : sp_level_of_count(dres[1]);
    int bmax = std::max(g.bdv[0], std::max(g.bdv[1], g.bdv[2]));
    int K = 0;
    while ((1 << K) < bmax) K++;
    const int GL = K + maxLevel;
    if (tern && (maxLevel < 1 || maxLevel > 6)) {
        set_error("shp_fmm_sph: translational maxLevel %d (the reference's ternary tree: 1..6; 0 is its "
                  "supercell-only path, 7 overflows its int box index)", maxLevel);
        return SHP_ERR_UNSUPPORTED;
    }
    if (!tern && (maxLevel < 2 || GL > 21)) {
        set_error("shp_fmm_sph: maxLevel %d with %d levels above the domains (need 2 <= maxLevel, total <= 21)",
                  maxLevel, K);
        return SHP_ERR_UNSUPPORTED;
    }
    // levels lmin .. GL carry expansions: the octree from 2, the ternary tree from 1
    const int lmin = tern ? 1 : 2;
    g.side = tern ? sp_pow3(maxLevel) : 1 << maxLevel;
    g.bsf = g.root / (float)g.side;
    g.bsd = tern ? (double)g.root / std::pow(3.0, maxLevel) : (double)g.root / g.side;
    g.GL = GL;
    // keys, sort, leaf boxes
    const dim3 gs(sp_bdiv(ns, 256)), bs(256);
    hipLaunchKernelGGL(k_sp_keys, gs, bs, 0, stream, P, sel, ns, g, kin, vin);
    size_t t1 = 0, t2 = 0, t3 = 0;
    const int kbits = tern ? 5 * GL : 3 * GL;  // 27^GL < 2^(5 GL)
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t1, kin, kout, vin, vout, ns, 0, kbits, stream));
    SHP_HIP_CHECK(hipcub::DeviceRunLengthEncode::Encode(nullptr, t2, kout, kin, vin, nruns, ns, stream));
    SHP_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t3, vin, vout, ns, stream));
    SHP_WS(cub2, char, h, WS_CUB, std::max(t1, std::max(t2, t3)));
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub2, t1, kin, kout, vin, vout, ns, 0, kbits, stream));
    // WS_SPH_PART: sorted sources (float4), leaf box of each, run flags/counts and starts
    char* pb = static_cast<char*>(h->get(WS_SPH_PART, (size_t)ns * (sizeof(float4) + 3 * sizeof(int)) + 64));
    if (!pb) return SHP_ERR_NOMEM;
    float4* sp4 = (float4*)pb;
    int* sbox = (int*)(sp4 + ns);
    int* lcnt = sbox + ns;
    int* lfirst = lcnt + ns;
    hipLaunchKernelGGL(k_sp_gather, gs, bs, 0, stream, P, vout, ns, kout, sp4, lfirst);
    SHP_HIP_CHECK(hipcub::DeviceScan::InclusiveSum(cub2, t3, lfirst, sbox, ns, stream));
    // leaf level: unique keys into the level store
    std::vector<long> loff(GL + 2, 0);
    std::vector<int> nbl(GL + 1, 0);
    // level keys: at most ns per level
    uint64_t* lk = static_cast<uint64_t*>(h->get(WS_SPH_LVL, sizeof(uint64_t) * (size_t)ns * (GL - 1) + 64));
    if (!lk) return SHP_ERR_NOMEM;
    SHP_HIP_CHECK(hipcub::DeviceRunLengthEncode::Encode(cub2, t2, kout, lk, lcnt, nruns, ns, stream));
    int nrun = 0;
    SHP_HIP_CHECK(hipMemcpyAsync(&nrun, nruns, sizeof(int), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    size_t t4 = 0;
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t4, lcnt, lfirst, nrun, stream));
    SHP_WS(cub3, char, h, WS_CUB, std::max(std::max(t1, t2), std::max(t3, t4)));
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub3, t4, lcnt, lfirst, nrun, stream));
    nbl[GL] = nrun;
    // level l occupies lk[loff[l] .. loff[l] + nbl[l]); leaf level first
    loff[GL] = 0;
    long used = nrun;
    for (int l = GL - 1; l >= lmin; l--) {
        uint64_t* child = lk + loff[l + 1];
        uint64_t* tmp = kin;  // free now
        hipLaunchKernelGGL(k_sp_shift, dim3(sp_bdiv(nbl[l + 1], 256)), dim3(256), 0, stream, child, nbl[l + 1], tern,
                           tmp);
        SHP_HIP_CHECK(hipcub::DeviceRunLengthEncode::Encode(cub3, t2, tmp, lk + used, vin, nruns, nbl[l + 1], stream));
        SHP_HIP_CHECK(hipMemcpyAsync(&nrun, nruns, sizeof(int), hipMemcpyDeviceToHost, stream));
        SHP_HIP_CHECK(hipStreamSynchronize(stream));
        loff[l] = used;
        nbl[l] = nrun;
        used += nrun;
    }
    // expansions: M and L of every level, KP doubles per box
    std::vector<long> eoff(GL + 1, 0);
    long etot = 0;
    for (int l = lmin; l <= GL; l++) {
        eoff[l] = etot;
        etot += (long)nbl[l] * KP;
    }
    double* E = static_cast<double*>(h->get(WS_SPH_EXP, sizeof(double) * 2 * (size_t)etot));
    if (!E) return SHP_ERR_NOMEM;
    double* Mx = E;
    double* Lx = E + etot;
    // operators: nch M2M, nch L2L (8 octants or 27 ternary children), the M2L
    // slots of one level (offsets in [-3, 3]^3 or [-5, 5]^3), 2 rotational M2M
    const int nch = tern ? 27 : 8;
    const int nslot = tern ? 1331 : kSpOps;
    const int nops = 2 * nch + nslot + 2;
    double* Tops = static_cast<double*>(h->get(WS_SPH_OPS, sizeof(double) * (size_t)nops * KP * KP));
    if (!Tops) return SHP_ERR_NOMEM;
    // descriptors and octant offsets (host, uploaded with the tiles)
    std::vector<int4> desc(nops);
    for (int o = 0; o < nch; o++) {
        desc[o] = make_int4(tern ? 4 : 0, o, 0, 0);
        desc[nch + o] = make_int4(tern ? 5 : 1, o, 0, 0);
    }
    const int sw = tern ? 11 : 7, sh = sw / 2;
    for (int s = 0; s < nslot; s++) {
        const int dx = s / (sw * sw) - sh, dy = (s / sw) % sw - sh, dz = s % sw - sh;
        const bool far = std::abs(dx) > 1 || std::abs(dy) > 1 || std::abs(dz) > 1;
        desc[2 * nch + s] = far ? make_int4(2, dx, dy, dz) : make_int4(-1, 0, 0, 0);
    }
    desc[2 * nch + nslot] = make_int4(3, 0, 0, 0);
    desc[2 * nch + nslot + 1] = make_int4(3, 1, 0, 0);
    // order m of each real component (the M2L gather's quadrant phase)
    std::vector<signed char> mord(KP, 0);
    for (int c = 0; c < nc2; c++) {
        const int ci = c >> 1;
        int n = 0;
        while ((n + 1) * (n + 2) / 2 <= ci) n++;
        mord[c] = (signed char)(ci - n * (n + 1) / 2);
    }
    std::vector<int4> offs;
    if (tern) sp_offsets_tern(offs);
    else sp_offsets(offs);
    const int noff = tern ? 702 : kSpNoff;
    // tiles: at most nb / 16 + one per child position per level
    int maxnb = 0;
    for (int l = lmin; l <= GL; l++) maxnb = std::max(maxnb, nbl[l]);
    const int maxtiles = maxnb / 16 + nch;
    char* gb = static_cast<char*>(h->get(WS_SPH_GRP, sizeof(int4) * ((size_t)nops + offs.size() + maxtiles) +
                                                          sizeof(int) * 4 * (size_t)maxnb + 256 + 128));
    if (!gb) return SHP_ERR_NOMEM;
    int4* ddesc = (int4*)gb;
    int4* doffs = ddesc + nops;
    int4* dtiles = doffs + offs.size();
    unsigned* okey = (unsigned*)(dtiles + maxtiles);
    unsigned* okey2 = okey + maxnb;
    int* oval = (int*)(okey2 + maxnb);
    int* grp = oval + maxnb;
    signed char* dmord = (signed char*)(grp + maxnb);
    SHP_HIP_CHECK(hipMemcpyAsync(ddesc, desc.data(), sizeof(int4) * nops, hipMemcpyHostToDevice, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(doffs, offs.data(), sizeof(int4) * offs.size(), hipMemcpyHostToDevice, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(dmord, mord.data(), (size_t)KP, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(k_sp_build_ops, dim3(2 * nch), dim3(128), 0, stream, S->tab, p, KP, ddesc, 1.0, 0, Tops);
    double* Trot = Tops + (size_t)(2 * nch + nslot) * KP * KP;
    // the four-image M2M into the domains (level K) is used when it is a level of the tree
    const bool rot_root = rot && K >= 2;
    if (rot_root)
        hipLaunchKernelGGL(k_sp_build_ops, dim3(2), dim3(128), 0, stream, S->tab, p, KP, ddesc + 2 * nch + nslot, 1.0,
                           0, Trot);
    // P2M, M2M
    hipLaunchKernelGGL(k_sp_p2m, dim3(nbl[GL]), dim3(64), 0, stream, S->tab, p, KP, g, lk + loff[GL], lfirst, lcnt,
                       sp4, Mx + eoff[GL]);
    for (int l = GL - 1; l >= lmin; l--) {
        const bool zop = rot_root && l == K;
        hipLaunchKernelGGL(k_sp_m2m, dim3(nbl[l]), dim3(128), 0, stream, zop ? Trot : Tops, KP, nc2, zop ? 1 : 0, tern,
                           lk + loff[l + 1], nbl[l + 1], Mx + eoff[l + 1], lk + loff[l], Mx + eoff[l]);
    }
    // downward pass
    double* Tm2l = Tops + (size_t)2 * nch * KP * KP;
    std::vector<int4> tiles;
    for (int l = lmin; l <= GL; l++) {
        const int nb = nbl[l];
        if (l == lmin) {
            hipLaunchKernelGGL(k_sp_zero, dim3(sp_bdiv((long)nb * KP, 256)), dim3(256), 0, stream, Lx + eoff[l],
                               (size_t)nb * KP);
        } else {
            hipLaunchKernelGGL(k_sp_l2l, dim3(nb), dim3(128), 0, stream, Tops + (size_t)nch * KP * KP, KP, nc2, tern,
                               lk + loff[l - 1], nbl[l - 1], Lx + eoff[l - 1], lk + loff[l], Lx + eoff[l]);
        }
        // box size at level l (m2l_kernel_translations: rootBoxSize / 3^l)
        const double hl = tern ? (double)g.root / std::pow(3.0, l) : (double)g.root * std::ldexp(1.0, K - l);
        hipLaunchKernelGGL(k_sp_build_ops, dim3(nslot), dim3(128), 0, stream, S->tab, p, KP, ddesc + 2 * nch, hl, 1,
                           Tm2l);
        // boxes grouped by child position (stable), tiles of 16
        hipLaunchKernelGGL(k_sp_zero_i, dim3(1), dim3(64), 0, stream, ocnt, (size_t)nch);
        hipLaunchKernelGGL(k_sp_octants, dim3(sp_bdiv(nb, 256)), dim3(256), 0, stream, lk + loff[l], nb, tern, okey,
                           oval, ocnt);
        const int obits = tern ? 5 : 3;
        size_t t5 = 0;
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t5, okey, okey2, oval, grp, nb, 0, obits, stream));
        SHP_WS(cub4, char, h, WS_CUB, t5);
        SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub4, t5, okey, okey2, oval, grp, nb, 0, obits, stream));
        int oc[27];
        SHP_HIP_CHECK(hipMemcpyAsync(oc, ocnt, sizeof(int) * nch, hipMemcpyDeviceToHost, stream));
        SHP_HIP_CHECK(hipStreamSynchronize(stream));
        tiles.clear();
        int start = 0;
        for (int o = 0; o < nch; o++) {
            for (int q = 0; q < oc[o]; q += 16) tiles.push_back(make_int4(o, start + q, std::min(16, oc[o] - q), 0));
            start += oc[o];
        }
        if (tiles.empty()) continue;
        SHP_HIP_CHECK(hipMemcpyAsync(dtiles, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice, stream));
        const int nt = (int)tiles.size();
        const uint64_t* keys = lk + loff[l];
        const int jd = (rot && l > K) ? (1 << (l - K)) : 0;  // the quadrant window below the domains
        // ternary levels of up to 243 cells per axis: a dense cell -> box table (<= 57 MB)
        int* cells = nullptr;
        if (tern && l <= 5) {
            const int W = sp_pow3(l);
            const size_t ncell = (size_t)W * W * W;
            cells = static_cast<int*>(h->get(WS_SPH_CELLS, sizeof(int) * ncell));
            if (!cells) return SHP_ERR_NOMEM;
            hipLaunchKernelGGL(k_sp_fill_i, dim3(sp_bdiv((long)ncell, 256)), dim3(256), 0, stream, cells, ncell, -1);
            hipLaunchKernelGGL(k_sp_cells_set, dim3(sp_bdiv(nb, 256)), dim3(256), 0, stream, keys, nb, W, cells);
        }
        // offset chunks per tile: about 8 waves per CU over the level, at most 27
        const int nck = std::max(1, std::min(27, (8 * h->num_cus + nt - 1) / nt));
        double* part = nullptr;
        if (nck > 1) {
            part = static_cast<double*>(h->get(WS_SPH_M2LP, sizeof(double) * (size_t)nck * nt * 16 * KP));
            if (!part) return SHP_ERR_NOMEM;
        }
        switch (RT) {
        case 1: sp_launch_m2l<1>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        case 2: sp_launch_m2l<2>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        case 3: sp_launch_m2l<3>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        case 4: sp_launch_m2l<4>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        case 5: sp_launch_m2l<5>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        case 6: sp_launch_m2l<6>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        default: sp_launch_m2l<7>(stream, nt, nck, part, Tm2l, Mx + eoff[l], Lx + eoff[l], keys, nb, l, jd, tern, noff, grp, dtiles, doffs, dmord, cells, pctr); break;
        }
        // the tile list is host memory reused by the next level: wait for the copy
        SHP_HIP_CHECK(hipStreamSynchronize(stream));
    }
    // leaves: P2P in units of 64 targets, longest first, then L2P
    const int nleaf = nbl[GL];
    // WS_SPH_NEAR: P2P sums [3][ns] doubles, then per leaf: unit count, offset,
    // cost; per unit (at most ns / 64 + nleaf): keys and values, in and out
    const size_t umax = (size_t)ns / 64 + nleaf + 1;
    char* nbuf = static_cast<char*>(h->get(WS_SPH_NEAR, sizeof(double) * 3 * (size_t)ns +
                                                             sizeof(unsigned long long) * 2 * umax +
                                                             sizeof(int) * (3 * (size_t)nleaf + 2 * umax) + 64));
    if (!nbuf) return SHP_ERR_NOMEM;
    double* near = (double*)nbuf;
    unsigned long long* uval = (unsigned long long*)(near + 3 * (size_t)ns);
    unsigned long long* uval2 = uval + umax;
    int* ucnt = (int*)(uval2 + umax);
    int* uoff = ucnt + nleaf;
    unsigned* ucost = (unsigned*)(uoff + nleaf);
    unsigned* ukey = ucost + nleaf;
    unsigned* ukey2 = ukey + umax;
    hipLaunchKernelGGL(k_sp_unit_counts, dim3(sp_bdiv(nleaf, 64)), dim3(64), 0, stream, g, lk + loff[GL], nleaf,
                       lcnt, ucnt, ucost, pctr);
    size_t t6 = 0, t7 = 0;
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, t6, ucnt, uoff, nleaf, stream));
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, t7, ukey, ukey2, uval, uval2, (int)umax, 0, 32, stream));
    SHP_WS(cub5, char, h, WS_CUB, std::max(t6, t7));
    SHP_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(cub5, t6, ucnt, uoff, nleaf, stream));
    int lastc[2];
    SHP_HIP_CHECK(hipMemcpyAsync(&lastc[0], uoff + nleaf - 1, sizeof(int), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipMemcpyAsync(&lastc[1], ucnt + nleaf - 1, sizeof(int), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    const int nunits = lastc[0] + lastc[1];
    hipLaunchKernelGGL(k_sp_unit_fill, dim3(sp_bdiv(nleaf, 64)), dim3(64), 0, stream, nleaf, ucnt, uoff, ucost, ukey,
                       uval);
    SHP_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(cub5, t7, ukey, ukey2, uval, uval2, nunits, 0, 32, stream));
    hipLaunchKernelGGL(k_sp_p2p, dim3(nunits), dim3(64), 0, stream, g, k_const_host(), lk + loff[GL], nleaf, lfirst,
                       lcnt, uval2, sp4, vout, ns, st, spc, near);
    hipLaunchKernelGGL(k_sp_leaf, dim3(sp_bdiv(ns, 64)), dim3(64), 0, stream, S->tab, p, KP, g, k_const_host(),
                       lk + loff[GL], sbox, sp4, vout, ns, st, spc, near, Lx + eoff[GL], i0, m, out);
    unsigned long long pc[3] = {0, 0, 0};
    SHP_HIP_CHECK(hipMemcpyAsync(pc, pctr, sizeof(pc), hipMemcpyDeviceToHost, stream));
    SHP_HIP_CHECK(hipStreamSynchronize(stream));
    long boxes = 0;
    for (int l = lmin; l <= GL; l++) boxes += nbl[l];
    double* sts = h->sph_stats;
    sts[0] = maxLevel;
    sts[1] = K;
    sts[2] = g.bdv[0];
    sts[3] = g.bdv[1];
    sts[4] = g.bdv[2];
    sts[5] = nbl[GL];
    sts[6] = (double)pc[0];
    sts[7] = (double)pc[1];
    sts[8] = ns;
    sts[9] = (double)boxes;
    sts[10] = (double)pc[2];
    sts[11] = 0;
    h->sph_have_stats = true;
    return SHP_OK;
}

}  // namespace

void sph_release(shp_handle h) {
    if (!h || !h->sph) return;
    SphState* S = static_cast<SphState*>(h->sph);
    if (S->buf) (void)hipFree(S->buf);
    delete S;
    h->sph = nullptr;
}

}  // namespace shp

using namespace shp;

extern "C" {

int shp_fmm_sph_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1,
                     int order, int max_level, const shp_symmetry* sym, double* out3) {
    SHP_REQUIRE(h && pos4 && vel4 && acc4 && out3, "shp_fmm_sph_host: null argument");
    SHP_REQUIRE(n >= 0 && 0 <= i0 && i0 <= i1 && i1 <= n, "shp_fmm_sph_host: bad range n=%d [%d,%d)", n, i0, i1);
    SHP_REQUIRE(order >= 1 && order <= kSpPmax, "shp_fmm_sph_host: order %d outside 1..%d", order, kSpPmax);
    SHP_REQUIRE(max_level == -2 || max_level >= 0, "shp_fmm_sph_host: max_level %d (-2: from the counts)", max_level);
    const int stype = sym ? sym->type : SHP_SYM_NONE;
    SHP_REQUIRE(stype == SHP_SYM_NONE || stype == SHP_SYM_ROTATIONAL || stype == SHP_SYM_TRANSLATIONAL,
                "shp_fmm_sph_host: symmetry type %d", stype);
    if (stype == SHP_SYM_ROTATIONAL && sym->nsec != 4) {
        set_error("shp_fmm_sph_host: the reference's rotational FMM has 4 sectors (got %d)", sym->nsec);
        return SHP_ERR_UNSUPPORTED;
    }
    float cell = 0.0f;
    if (stype == SHP_SYM_TRANSLATIONAL) {
        // the reference's translational FMM: one cubic periodic cell on all three
        // axes (rootBoxSize asserted equal on the periodic axes, domain_bounds_
        // functions.cpp:600-610) and the 27 images of shifts {-1, 0, 1}^3
        cell = (float)sym->domain[0];
        bool ok = cell > 0 && (float)sym->domain[1] == cell && (float)sym->domain[2] == cell && sym->nimg == 27 &&
                  sym->lattice != nullptr;
        int seen[27] = {0};
        for (int k = 0; ok && k < 27; k++) {
            const int* q = sym->lattice + 3 * k;
            if (std::abs(q[0]) > 1 || std::abs(q[1]) > 1 || std::abs(q[2]) > 1) ok = false;
            else if (seen[(q[0] + 1) * 9 + (q[1] + 1) * 3 + q[2] + 1]++) ok = false;
        }
        if (!ok) {
            set_error("shp_fmm_sph_host: the reference's translational FMM needs a cubic cell on all three axes and "
                      "the 27 lattice points {-1, 0, 1}^3 (got %d images)", sym->nimg);
            return SHP_ERR_UNSUPPORTED;
        }
    }
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
    hipLaunchKernelGGL(k_sp_from_aos, dim3(sp_bdiv(n, 256)), dim3(256), 0, h->stream, dvel, dacc, n, ints, ints + n);
    int rc = sph_eval(h, dpos, ints, ints + n, i0, i1, order, max_level, stype == SHP_SYM_ROTATIONAL ? 1 : 0, cell,
                      dout);
    if (rc) return rc;
    hipLaunchKernelGGL(k_sp_interleave, dim3(sp_bdiv(m, 256)), dim3(256), 0, h->stream, dout, m, dout + 3 * (size_t)m);
    SHP_HIP_CHECK(hipMemcpyAsync(out3, dout + 3 * (size_t)m, sizeof(double) * 3 * (size_t)m, hipMemcpyDeviceToHost,
                                 h->stream));
    SHP_HIP_CHECK(hipStreamSynchronize(h->stream));
    return SHP_OK;
}

int shp_fmm_sph_stats(shp_handle h, double* out) {
    SHP_REQUIRE(h && out, "shp_fmm_sph_stats: null argument");
    SHP_REQUIRE(h->sph_have_stats, "shp_fmm_sph_stats: no spherical FMM evaluation on this handle yet");
    std::memcpy(out, h->sph_stats, sizeof(h->sph_stats));
    return SHP_OK;
}

}  // extern "C"
