/*
 * oracle/fmm_sph.c -- CPU FP64 restatement of the reference's FMM operators:
 * spherical-harmonic expansions of 1/r translated by rotation, z-axis shift
 * and rotation back.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/direct.c).
 *
 * Follows, operation for operation where the rounding is visible:
 *   tables      MultipoleMethodConstants::precalc   multipole_method_constants.cpp:40-226
 *               (factorial, anm, Anm stored in float; Ynm, Dnm complex double)
 *   rotation    MultipoleMethodForceCalculation::rotation   fmm_cpukernel.cpp:10-32
 *   P2M         p2m_kernel   fmm_cpukernel.cpp:281-384
 *   M2M         m2m_kernel   fmm_cpukernel.cpp:513-619
 *   M2L         m2l_kernel   fmm_cpukernel.cpp:742-909 (softened centre distance :832,
 *               direction reduced by gcd :836-872)
 *   L2L         l2l_kernel   fmm_cpukernel.cpp:1050-1126
 *   L2P         l2p_kernel   fmm_cpukernel.cpp:1188-1317
 *   P2P         p2p_kernel   fmm_cpukernel.cpp:194-219 over DirectForceFunction::direct
 *               (direct_force_method.cpp:153-190)
 *   cart2sph_exact  auxiliary_functions.cpp:160-186 (float sqrtf / acosf / atanf)
 *   rotation index  rotationMorton multipole_method_force_calculation.cpp:1610-1652,
 *               unmorton11 multipole_method_functions.cpp
 *
 * Tree.  setDomainSize (multipole_method_force_calculation.cpp:648-832) lays a
 * grid of cubic domains of side rootBoxSize from computeDomainBounds; every
 * domain holds an octree of depth maxLevel (setOptimumLevel :835-881, from the
 * largest domain's particle count, or setMaxLevel).  Particles are placed in
 * their domain and then in their leaf with the reference's float arithmetic
 * (setDomainSize :714-770, morton :976-1057).  Below the domain level the
 * reference's P2P and M2L lists (getInteractionList :1685-2446) work in global
 * cell coordinates across domains, which is the octree of one cube of side
 * 2^K rootBoxSize (2^K >= the largest domain count) cut K levels above the
 * domains.  This file uses that octree at every level, so for a single domain
 * (K = 0) the tree, the lists and their order are the reference's.  With
 * several domains the reference groups the domains into ternary supercells
 * above the domain level (sm2m / sm2l); here those levels are octree levels
 * of the enclosing cube.  Levels: P2M at the leaves, M2M up to level 2, M2L at
 * levels 2 .. leaf (children of the parent's 27 neighbours that are not
 * neighbours), L2L down, L2P + 27-neighbour P2P at the leaves.
 *
 * Rotational symmetry (updateAccelerationRotational, multipole_method_force_
 * calculation_rotational.cpp:9-331): the ions fill one quadrant (x >= 0,
 * y >= 0) of a domain column centred on the z axis (computeDomainBoundsRotations,
 * domain_bounds_functions.cpp:317-500), and the other three quadrants hold
 * its images rotated by k pi / 2, k = 1, 2, 3.  Below the domain level the
 * P2P and M2L lists (getInteractionListRotations :568-1438) map a cell of
 * quadrant k back to the real cell it images, with k: P2P rotates that cell's
 * sources (DirectForceFunction::direct_rotational, direct_force_method_
 * rotational.cpp:48-94), M2L multiplies its multipole by i^(-m k) (m2l_kernel_
 * rotations, fmm_cpukernel_rotational.cpp:298-496, qrf_complex), and the M2M
 * into a domain adds the four images of each child (m2m_kernel_rotations :107-
 * 290).  Above the domain level there is no quadrant map ("supercells don't
 * have rotational symmetry").  The quadrant map works in domain-local x, y,
 * so the reference's lists are consistent only for one domain per x and y;
 * other grids are not restated (-4).
 *
 * Translational symmetry (orc_fmm_sph_trans; updateAccelerationTranslational,
 * multipole_method_force_calculation_translational.cpp:9-331): one cubic cell
 * periodic on all three axes, a ternary tree with base-27 keys, lists that
 * wrap periodically, the ternary M2M / L2L shifts.  Pinned by convergence to
 * the direct sum over the 27 images and by a brute-force near field; the
 * supercell passes of several domains along a non-periodic axis are not
 * restated.
 *
 * Expansion storage: coefficient (n, m), 0 <= m <= n < p, at n(n+1)/2 + m,
 * complex double.  Morton bit order of the reference (multipole_method_
 * functions.cpp unmorton): y -> bit 3i, x -> bit 3i+1, z -> bit 3i+2.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_domain_bounds(const float* pos4, const float* acc4, int i0, int i1, const int* forced3, int rot,
                      float* box_min3, float* dims3, int* nd3);
double orc_k_const(void);
long double orc_species_mass_ld(int s);

#define SP_PMAX 10                            /* FMMConstant::maxNumExpansions */
#define SP_NE2 (SP_PMAX * SP_PMAX)            /* numExpansion2 */
#define SP_NE4 (SP_NE2 * SP_NE2)              /* numExpansion4 */
#define SP_NCOEF (SP_PMAX * (SP_PMAX + 1) / 2) /* numCoefficients */
#define SP_NROT 1155                          /* possibleRotations */
#define SP_NREL 1331                          /* numRelativeBox = 11^3 */
#define SP_SOFTENING 6.4e-19                  /* PhysicalConstant::SOFTENING */
static const long double SP_PI = 3.14159265358979323846264338328L; /* PhysicalConstant::PI (long double) */

typedef struct { double re, im; } cz;

static inline cz cz_make(double r, double i) { cz c = {r, i}; return c; }
static inline cz cz_add(cz a, cz b) { return cz_make(a.re + b.re, a.im + b.im); }
static inline cz cz_mul(cz a, cz b) { return cz_make(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cz cz_scale(cz a, double s) { return cz_make(a.re * s, a.im * s); }
static inline cz cz_conj(cz a) { return cz_make(a.re, -a.im); }
/* std::exp(complex(0, t)) */
static inline cz cz_expi(double t) { return cz_make(cos(t), sin(t)); }

/* ------------------------------------------------------------------ tables -- */

static float g_factorial[4 * SP_NE2];
static float g_anm[4 * SP_NE2];
static float g_Anm[SP_NE4];
static cz g_Ynm[4 * SP_NE2];
static cz g_Dnm[SP_NROT][SP_PMAX][SP_NE2];
static int g_rotationLink[SP_NREL];
static int g_tables_ready = 0;

static int gcd2(int a, int b) {
    a = abs(a);
    b = abs(b);
    while (b != 0) {
        const int t = b;
        b = a % b;
        a = t;
    }
    return a;
}
static int gcd3(int a, int b, int c) { return gcd2(a, gcd2(b, c)); }

/* Float arc functions of cart2sph_exact.  By default this host's libm
 * acosf / atanf, as the reference calls them; with orc_sph_set_cr_trig(1) the
 * correctly rounded float values (the double function rounded to float),
 * which any libm that rounds correctly gives.  The two differ in the last
 * float bit for some arguments: the reference's results depend on its libm
 * there. */
static int g_cr_trig = 0;
static inline float sp_acosf(float x) { return g_cr_trig ? (float)acos((double)x) : acosf(x); }
static inline float sp_atanf(float x) { return g_cr_trig ? (float)atan((double)x) : atanf(x); }
void orc_sph_set_cr_trig(int on) {
    if ((on != 0) != g_cr_trig) g_tables_ready = 0;
    g_cr_trig = on != 0;
}

/* cart2sph_exact (auxiliary_functions.cpp:160-186): float libm on doubles. */
static void cart2sph_exact(double* r, double* theta, double* phi, double dx, double dy, double dz) {
    *r = sqrtf((float)(dx * dx + dy * dy + dz * dz));
    if (*r != 0) {
        *theta = sp_acosf((float)(dz / *r));
        if (*theta == 0 && dz != 0) *theta = sqrtf((float)(dx * dx + dy * dy)) / dz;
    } else {
        *theta = 0;
    }
    if (dx == 0 && dy == 0) *phi = 0;
    else if (dx == 0) *phi = (double)(dy / fabs(dy) * SP_PI * 0.5);
    else if (dx > 0) *phi = sp_atanf((float)(dy / dx));
    else *phi = (double)(sp_atanf((float)(dy / dx)) + SP_PI);
}

/* unmorton11 (multipole_method_functions.cpp): base-11 digits y, x, z. */
static void unmorton11(int i, int* x, int* y, int* z) {
    int m3[3] = {0, 0, 0};
    int n = i, k = 0, lvl = 0, pw = 1;
    while (n != 0) {
        const int j = 2 - k;
        if (lvl == 0) m3[j] += n % 11;
        else m3[j] += (n % 11) * pw;
        n /= 11;
        k = (k + 1) % 3;
        if (k == 0) { pw *= 11; lvl++; }
    }
    *x = m3[1];
    *y = m3[2];
    *z = m3[0];
}

#define NMK(n, m, k) ((4 * (n) * (n) * (n) + 6 * (n) * (n) + 5 * (n)) / 3 + (m) * (2 * (n) + 1) + (k))

/* MultipoleMethodConstants::precalc (multipole_method_constants.cpp:40-226). */
static void sph_precalc(void) {
    if (g_tables_ready) return;
    static double anmk[SP_NE4][2];
    static double Dnmd[SP_NE4];
    const int P = SP_PMAX;
    for (int n = 0; n < 2 * P; n++) {
        for (int m = -n; m <= n; m++) {
            const int nm = n * n + n + m, nabsm = abs(m);
            double fnmm = 1.0, fnpm = 1.0, fnma = 1.0, fnpa = 1.0;
            for (int i = 1; i <= n - m; i++) fnmm *= i;
            for (int i = 1; i <= n + m; i++) fnpm *= i;
            for (int i = 1; i <= n - nabsm; i++) fnma *= i;
            for (int i = 1; i <= n + nabsm; i++) fnpa *= i;
            g_factorial[nm] = (float)sqrt(fnma / fnpa);
            const double fad = sqrt(fnmm * fnpm);
            g_anm[nm] = (float)(pow(-1.0, n) / fad);
        }
    }
    for (int j = 0; j < P; j++) {
        for (int k = -j; k <= j; k++) {
            const int jk = j * j + j + k;
            for (int n = abs(k); n < P; n++) {
                const int nk = n * n + n + k;
                const int jkn = jk * SP_NE2 + nk;
                const int jnk = (j + n) * (j + n) + j + n;
                g_Anm[jkn] = (float)(pow(-1.0, j + k) * g_anm[nk] * g_anm[jk] / g_anm[jnk]);
            }
        }
    }
    double pn = 1;
    for (int m = 0; m < 2 * P; m++) {
        double p = pn;
        const int npn = m * m + 2 * m, nmn = m * m;
        g_Ynm[npn] = cz_make(g_factorial[npn] * p, 0.0);
        g_Ynm[nmn] = cz_conj(g_Ynm[npn]);
        double p1 = p;
        p = (2 * m + 1) * p;
        for (int n = m + 1; n < 2 * P; n++) {
            const int npm = n * n + n + m, nmm = n * n + n - m;
            g_Ynm[npm] = cz_make(g_factorial[npm] * p, 0.0);
            g_Ynm[nmm] = cz_conj(g_Ynm[npm]);
            const double p2 = p1;
            p1 = p;
            p = ((2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
        }
        pn = 0;
    }
    for (int n = 0; n < P; n++) {
        for (int m = 1; m <= n; m++) {
            const double anmd = n * (n + 1) - m * (m - 1);
            for (int k = 1 - m; k < m; k++) {
                const int nmk = NMK(n, m, k);
                const double anmkd = ((double)(n * (n + 1) - k * (k + 1))) / (n * (n + 1) - m * (m - 1));
                anmk[nmk][0] = -(m + k) / sqrt(anmd);
                anmk[nmk][1] = sqrt(anmkd);
            }
        }
    }
    int nrot = 0;
    for (int i = 0; i < SP_NREL; i++) {
        int bx, by, bz;
        unmorton11(i, &bx, &by, &bz);
        const double dx = bx - 5, dy = by - 5, dz = bz - 5;
        if (gcd3(bx - 5, by - 5, bz - 5) > 1) {
            g_rotationLink[i] = -1;
            continue;
        }
        g_rotationLink[i] = nrot;
        double rho, alpha, beta;
        cart2sph_exact(&rho, &alpha, &beta, dx, dy, dz);
        if (rho == 0) alpha = 0;
        const double sc = sin(alpha) / (1 + cos(alpha));
        cz expBeta[4 * SP_PMAX];
        for (int n = 0; n < 4 * P - 3; n++) expBeta[n] = cz_expi((n - 2 * P + 2) * beta);
        const int south = (rho == -dz) && (rho > 0);
        for (int n = 0; n < P; n++) {
            int nmk = NMK(n, n, n);
            Dnmd[nmk] = pow(cos(alpha * 0.5), 2 * n);
            if (south && n > 0) Dnmd[nmk] = 0;
            for (int k = n; k >= 1 - n; k--) {
                nmk = NMK(n, n, k);
                const int nmk1 = NMK(n, n, k - 1);
                const double ank = ((double)n + k) / (n - k + 1);
                Dnmd[nmk1] = sqrt(ank) * tan(alpha * 0.5) * Dnmd[nmk];
                if (south) {
                    Dnmd[nmk1] = 0;
                    if (k == 1 - n) Dnmd[nmk1] = 1;
                }
            }
            int count = 0;
            for (int m = n; m >= 1; m--) {
                for (int k = m - 1; k >= 1 - m; k--) {
                    const int a = NMK(n, m, k), a1 = NMK(n, m, k + 1), m1 = NMK(n, m - 1, k);
                    Dnmd[m1] = anmk[a][1] * Dnmd[a1] + anmk[a][0] * sc * Dnmd[a];
                    if (south) {
                        Dnmd[m1] = 0;
                        count++;
                        if (k == 1 - m) Dnmd[m1] = pow(-1.0, count);
                    }
                }
            }
        }
        for (int n = 1; n < P; n++) {
            for (int m = 0; m <= n; m++) {
                for (int k = -m; k <= -1; k++) {
                    const double ek = pow(-1.0, k);
                    const int a = NMK(n, m, k);
                    const int b = (4 * n * n * n + 6 * n * n + 5 * n) / 3 - k * (2 * n + 1) - m;
                    Dnmd[a] = ek * Dnmd[a];
                    Dnmd[b] = pow(-1.0, m + k) * Dnmd[a];
                }
                for (int k = 0; k <= m; k++) {
                    const int a = NMK(n, m, k);
                    const int b = NMK(n, k, m);
                    const int c = (4 * n * n * n + 6 * n * n + 5 * n) / 3 - k * (2 * n + 1) - m;
                    Dnmd[b] = pow(-1.0, m + k) * Dnmd[a];
                    Dnmd[c] = Dnmd[b];
                }
            }
        }
        for (int n = 0; n < P; n++)
            for (int m = 0; m <= n; m++)
                for (int k = -n; k <= n; k++) {
                    const int nk = n * (n + 1) + k;
                    g_Dnm[nrot][m][nk] = cz_scale(expBeta[k + m + 2 * P - 2], Dnmd[NMK(n, m, k)]);
                }
        nrot++;
    }
    g_tables_ready = 1;
}

/* rotationMorton with an 11-grid index (multipole_method_force_calculation.cpp:1610-1652):
 * the forward rotation of direction (x, y, z) in [0, 11)^3 and the backward one of
 * (10 - x, y, z), both through rotationLink. */
static void rotation_index(int x, int y, int z, int* fwd, int* bwd) {
    const int base = y + z * 121;
    *fwd = g_rotationLink[base + x * 11];
    *bwd = g_rotationLink[base + (10 - x) * 11];
}

/* MultipoleMethodForceCalculation::rotation (fmm_cpukernel.cpp:10-32). */
static void sph_rotate(int p, const cz* C, cz* out, int rot) {
    for (int n = 0; n < p; n++) {
        for (int m = 0; m <= n; m++) {
            cz s = cz_make(0, 0);
            for (int k = -n; k <= -1; k++) s = cz_add(s, cz_mul(g_Dnm[rot][m][n * (n + 1) + k], cz_conj(C[n * (n + 1) / 2 - k])));
            for (int k = 0; k <= n; k++) s = cz_add(s, cz_mul(g_Dnm[rot][m][n * (n + 1) + k], C[n * (n + 1) / 2 + k]));
            out[n * (n + 1) / 2 + m] = s;
        }
    }
}

/* ------------------------------------------------------------- operators -- */

/* P2M of one particle (fmm_cpukernel.cpp:338-374): dist is the float difference
 * position - boxCenter widened to double, box_size the leaf size in double. */
static void sph_p2m_one(int p, double q, double dx, double dy, double dz, double box_size, cz* M) {
    double rho, alpha, beta, Y[SP_NE2];
    cart2sph_exact(&rho, &alpha, &beta, dx, dy, dz);
    rho /= box_size;
    const double xx = cos(alpha);
    const double s2 = sqrt((1 - xx) * (1 + xx));
    double fact = 1, pn = 1, rhom = 1.0;
    for (int m = 0; m < p; m++) {
        double pv = pn;
        int nm = m * m + 2 * m;
        Y[nm] = rhom * g_factorial[nm] * pv;
        double p1 = pv;
        pv = xx * (2 * m + 1) * pv;
        rhom *= rho;
        double rhon = rhom;
        for (int n = m + 1; n < p; n++) {
            nm = n * n + n + m;
            Y[nm] = rhon * g_factorial[nm] * pv;
            const double p2 = p1;
            p1 = pv;
            pv = (xx * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
            rhon *= rho;
        }
        pn = -pn * fact * s2;
        fact += 2;
    }
    for (int n = 0; n < p; n++)
        for (int m = 0; m <= n; m++) {
            const cz eim = cz_expi(-m * beta);
            const cz qy = cz_make(q * Y[n * n + n + m], 0.0 * Y[n * n + n + m]);
            M[n * (n + 1) / 2 + m] = cz_add(M[n * (n + 1) / 2 + m], cz_mul(qy, eim));
        }
}

/* M2M of one child into its parent: rotation by the 11-grid direction
 * (rx, ry, rz), the z shift by rho child sizes scaled base^-j, the rotation
 * back, accumulated (m2m_kernel, fmm_cpukernel.cpp:562-605: octree,
 * rho = sqrt(3) / 2, base 2; m2m_kernel_translations,
 * fmm_cpukernel_translational.cpp:507-547: ternary, rho = sqrt(|x-1| + |y-1| +
 * |z-1|), base 3). */
static void sph_m2m_g(int p, int rx, int ry, int rz, double rho, double base, const cz* Mc, cz* Mp) {
    int fwd, bwd;
    rotation_index(rx, ry, rz, &fwd, &bwd);
    cz A[SP_NCOEF], B[SP_NCOEF];
    sph_rotate(p, Mc, B, fwd);
    for (int j = 0; j < p; j++) {
        const double powB = pow(base, -j);
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            cz s = cz_make(0, 0);
            for (int n = 0; n <= j - abs(k); n++) {
                const int jnk = (j - n) * (j - n) + j - n + k;
                const int jnks = (j - n) * (j - n + 1) / 2 + k;
                const int nm = n * n + n;
                const double c0 = pow(-1.0, n) * g_anm[nm] * g_anm[jnk] / g_anm[jk] * pow(rho, n);
                const cz cnm = cz_scale(cz_scale(g_Ynm[nm], c0), powB);
                s = cz_add(s, cz_mul(B[jnks], cnm));
            }
            A[j * (j + 1) / 2 + k] = s;
        }
    }
    sph_rotate(p, A, B, bwd);
    for (int j = 0; j < p * (p + 1) / 2; j++) Mp[j] = cz_add(Mp[j], B[j]);
}

/* M2M of one child of octant (cx, cy, cz) into its parent (fmm_cpukernel.cpp:562-605). */
static void sph_m2m(int p, int cx, int cy, int cz_, const cz* Mc, cz* Mp) {
    sph_m2m_g(p, 6 - 2 * cx, 6 - 2 * cy, 6 - 2 * cz_, sqrt(3.0) / 2, 2.0, Mc, Mp);
}

/* M2L of one source box at integer offset (dx, dy, dz) = target - source, in
 * boxes of size h (fmm_cpukernel.cpp:810-905). */
static void sph_m2l(int p, int dx, int dy, int dz, double box_size, const cz* Ms, cz* Lt) {
    const double fx = dx, fy = dy, fz = dz;
    const double rho = sqrt(fx * fx + fy * fy + fz * fz + SP_SOFTENING / (box_size * box_size));
    const int g = gcd3(dx, dy, dz);
    switch (g) {
    case 2: dx >>= 1; dy >>= 1; dz >>= 1; break;
    case 3:
    case 5:
        dx = (dx > 0) - (dx < 0);
        dy = (dy > 0) - (dy < 0);
        dz = (dz > 0) - (dz < 0);
        break;
    case 4: dx >>= 2; dy >>= 2; dz >>= 2; break;
    default:
        if (g != 0) { dx /= g; dy /= g; dz /= g; }
        break;
    }
    int fwd, bwd;
    rotation_index(dx + 5, dy + 5, dz + 5, &fwd, &bwd);
    cz A[SP_NCOEF], LA[SP_NCOEF], LB[SP_NCOEF];
    sph_rotate(p, Ms, A, fwd);
    double rhoj = 1;
    for (int j = 0; j < p; j++) {
        double rhojk = rhoj;
        rhoj *= rho;
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            cz s = cz_make(0, 0);
            double rhojn = rhojk;
            rhojk *= rho;
            for (int n = abs(k); n < p; n++) {
                rhojn *= rho;
                const int nk = n * n + n + k;
                const int jkn = jk * SP_NE2 + nk;
                const int jnk = (j + n) * (j + n) + j + n;
                const cz cnm = cz_scale(g_Ynm[jnk], (double)g_Anm[jkn] / rhojn);
                s = cz_add(s, cz_mul(A[n * (n + 1) / 2 + k], cnm));
            }
            LA[j * (j + 1) / 2 + k] = s;
        }
    }
    sph_rotate(p, LA, LB, bwd);
    for (int j = 0; j < p * (p + 1) / 2; j++) Lt[j] = cz_add(Lt[j], LB[j]);
}

/* L2L from the parent into one child: rotation by the 11-grid direction
 * (rx, ry, rz), the z shift by rho child sizes scaled base^(-n-1), the
 * rotation back; the child's expansion is overwritten (l2l_kernel,
 * fmm_cpukernel.cpp:1096-1125: octree; l2l_kernel_translations,
 * fmm_cpukernel_translational.cpp:795-833: ternary). */
static void sph_l2l_g(int p, int rx, int ry, int rz, double rho, double base, const cz* Lp, cz* Lc) {
    int fwd, bwd;
    rotation_index(rx, ry, rz, &fwd, &bwd);
    cz A[SP_NCOEF], B[SP_NCOEF];
    sph_rotate(p, Lp, B, fwd);
    for (int j = 0; j < p; j++) {
        for (int k = 0; k <= j; k++) {
            const int jk = j * j + j + k;
            cz s = cz_make(0, 0);
            for (int n = j; n < p; n++) {
                const int jnk = (n - j) * (n - j) + n - j;
                const int nk = n * n + n + k;
                /* anm[jnk] * anm[jk] / anm[nk] is float arithmetic in the reference */
                const float f = g_anm[jnk] * g_anm[jk] / g_anm[nk];
                const cz cnm = cz_scale(cz_scale(g_Ynm[jnk], (double)f * pow(rho, n - j)), pow(base, -n - 1));
                s = cz_add(s, cz_mul(B[n * (n + 1) / 2 + k], cnm));
            }
            A[j * (j + 1) / 2 + k] = s;
        }
    }
    sph_rotate(p, A, B, bwd);
    memcpy(Lc, B, sizeof(cz) * (size_t)(p * (p + 1) / 2));
}

/* L2L from the parent into its child of octant (cx, cy, cz) (fmm_cpukernel.cpp:1096-1125). */
static void sph_l2l(int p, int cx, int cy, int cz_, const cz* Lp, cz* Lc) {
    sph_l2l_g(p, 2 * cx + 4, 2 * cy + 4, 2 * cz_ + 4, sqrt(3.0) / 2, 2.0, Lp, Lc);
}

/* L2P of one particle (fmm_cpukernel.cpp:1243-1303): the field vector accel
 * (before -K q / m) at dist from the box centre. */
static void sph_l2p_one(int p, const cz* L, double dx, double dy, double dz, double box_size, double acc[3]) {
    double r, theta, phi, Y[SP_NE2], Yt[SP_NE2];
    cart2sph_exact(&r, &theta, &phi, dx, dy, dz);
    const double xx = cos(theta), yy = sin(theta);
    const double s2 = sqrt((1 - xx) * (1 + xx));
    double fact = 1, pn = 1;
    for (int m = 0; m < p; m++) {
        double pv = pn;
        int nm = m * m + 2 * m;
        Y[nm] = g_factorial[nm] * pv;
        double p1 = pv;
        pv = xx * (2 * m + 1) * pv;
        Yt[nm] = g_factorial[nm] * (pv - (m + 1) * xx * p1) / yy;
        for (int n = m + 1; n < p; n++) {
            nm = n * n + n + m;
            Y[nm] = g_factorial[nm] * pv;
            const double p2 = p1;
            p1 = pv;
            pv = (xx * (2 * n + 1) * p1 - (n + m) * p2) / (n - m + 1);
            Yt[nm] = g_factorial[nm] * ((n - m + 1) * pv - (n + 1) * xx * p1) / yy;
        }
        pn = -pn * fact * s2;
        fact += 2;
    }
    double aR = 0, aT = 0, aP = 0, rn = 1;
    r /= box_size;
    for (int n = 0; n < p; n++) {
        int nm = n * n + n, nms = n * (n + 1) / 2;
        double rr = n * rn / r * Y[nm];
        double rt = rn * Yt[nm];
        aR += rr * L[nms].re;
        aT += rt * L[nms].re;
        for (int m = 1; m <= n; m++) {
            nm = n * n + n + m;
            nms = n * (n + 1) / 2 + m;
            const cz eim = cz_expi(m * phi);
            const cz crr = cz_scale(eim, n * rn / r * Y[nm]);
            const cz crt = cz_scale(eim, rn * Yt[nm]);
            const cz cm = cz_scale(eim, m * rn * Y[nm]);
            const cz crp = cz_make(-cm.im, cm.re); /* times I */
            aR += 2 * cz_mul(crr, L[nms]).re;
            aT += 2 * cz_mul(crt, L[nms]).re;
            aP += 2 * cz_mul(crp, L[nms]).re;
        }
        rn *= r;
    }
    aR /= (box_size * box_size);
    aT /= (box_size * box_size);
    aP /= (box_size * box_size);
    acc[0] = sin(theta) * cos(phi) * aR + cos(theta) * cos(phi) / r * aT - sin(phi) / r / sin(theta) * aP;
    acc[1] = sin(theta) * sin(phi) * aR + cos(theta) * sin(phi) / r * aT + cos(phi) / r / sin(theta) * aP;
    acc[2] = cos(theta) * aR - sin(theta) / r * aT;
}

/* ------------------------------------------------------------------- tree -- */

/* Morton interleave of a cell in the reference's bit order: y -> 3i, x -> 3i+1, z -> 3i+2. */
static uint64_t sp_key(uint32_t x, uint32_t y, uint32_t z, int levels) {
    uint64_t k = 0;
    for (int i = 0; i < levels; i++) {
        k |= (uint64_t)((y >> i) & 1u) << (3 * i);
        k |= (uint64_t)((x >> i) & 1u) << (3 * i + 1);
        k |= (uint64_t)((z >> i) & 1u) << (3 * i + 2);
    }
    return k;
}
static void sp_unkey(uint64_t k, int levels, int* x, int* y, int* z) {
    int a = 0, b = 0, c = 0;
    for (int i = 0; i < levels; i++) {
        b |= (int)((k >> (3 * i)) & 1u) << i;
        a |= (int)((k >> (3 * i + 1)) & 1u) << i;
        c |= (int)((k >> (3 * i + 2)) & 1u) << i;
    }
    *x = a;
    *y = b;
    *z = c;
}

typedef struct {
    int nb;         /* boxes at this level */
    uint64_t* key;  /* sorted */
    int* first;     /* particle runs (leaf level only) */
    int* count;
    cz* M;          /* nb * ncoef */
    cz* L;
} sp_level;

static int sp_find(const sp_level* lv, uint64_t key) {
    int lo = 0, hi = lv->nb - 1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (lv->key[mid] == key) return mid;
        if (lv->key[mid] < key) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

typedef struct { uint64_t key; int idx; } sp_ki;
static int sp_cmp(const void* a, const void* b) {
    const sp_ki* x = (const sp_ki*)a;
    const sp_ki* y = (const sp_ki*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx - y->idx;
}

static int sp_ilog2ceil(int v) {
    int k = 0;
    while ((1 << k) < v) k++;
    return k;
}

/* The quadrant map of getInteractionListRotations (multipole_method_force_
 * calculation_rotational.cpp:616-634): a cell (x, y) of the local window of
 * jd cells per axis lies in quadrant k of the axis at jd / 2 and images the
 * real cell (*rx, *ry) of quadrant 0 rotated by k pi / 2. */
static int sp_quadrant(int x, int y, int jd, int* rx, int* ry) {
    const int half = jd >> 1;
    if (x >= half && y >= half) {
        *rx = x;
        *ry = y;
        return 0;
    }
    if (x < half && y >= half) {
        *rx = y;
        *ry = jd - x - 1;
        return 1;
    }
    if (x < half && y < half) {
        *rx = jd - x - 1;
        *ry = jd - y - 1;
        return 2;
    }
    *rx = jd - y - 1;
    *ry = x;
    return 3;
}

/* quadrantRotationFunction (auxiliary_functions.h:55-57): sin(n pi / 2). */
static int sp_qrf(int n) {
    const int odd = (n & 1) == 1 ? -1 : 1;
    const int sgn = ((1 - odd) > 0) - ((1 - odd) < 0);
    return sgn * ((((n - 1) >> 1) & 1) == 1 ? -1 : 1);
}

/* A multipole imaged into quadrant k: C_n^m qrf_complex(-m k) (fmm_cpukernel_
 * rotational.cpp:367-372, :190-195). */
static void sp_phase(int p, int k, const cz* C, cz* out) {
    for (int n = 0; n < p; n++)
        for (int m = 0; m <= n; m++) {
            const cz r = cz_make((double)sp_qrf(-m * k + 1), (double)sp_qrf(-m * k));
            out[n * (n + 1) / 2 + m] = cz_mul(C[n * (n + 1) / 2 + m], r);
        }
}

/* setOptimumLevel's level_switch (multipole_method_force_calculation.cpp:842-869). */
static int sp_level_of_count(int c) {
    static const float sw[8] = {200.0f, 1000.0f, 1.3e4f, 1e5f, 7e5f, 5e6f, 3e7f, 1.5e8f};
    for (int i = 0; i < 8; i++)
        if ((float)c < sw[i]) return i;
    return 8;
}

/*
 * The reference's FMM (operators restated above) over the tree described at
 * the top, on the reference's float AoS.  Targets: status 1 in [i0, i1);
 * sources: status 1 or 2 in [i0, i1).  p: numExpansions (1..10).
 * forced_level: setMaxLevel's value (-2: setOptimumLevel).  sym: 0 none,
 * 1 rotational (the reference's four quadrants).
 * out3[3 (i - i0) + c]: the acceleration in double (zero for non-targets),
 * the reference's double terms before their float accumulation.
 * info (may be NULL, 12 entries): [0] maxLevel, [1] K, [2..4] boxDimVals,
 * [5] leaf boxes, [6] M2L box pairs, [7] P2P box pairs, [8] sources,
 * [9] boxes over all levels, [10] P2P ion pairs (leaf ions x neighbour-list ions).
 * Returns 0; -1 for p outside 1..10; -2 if a source lies outside the domain
 * grid (the reference then runs its outlier passes, not restated here);
 * -3 for maxLevel < 2 (the reference's single-box and supercell-only paths);
 * -4 for a rotational grid with more than one domain in x or y.
 */
int orc_fmm_sph(const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1, int p,
                int forced_level, int sym, double* out3, long long* info) {
    if (p < 1 || p > SP_PMAX || sym < 0 || sym > 1) return -1;
    const int rot = sym == 1;
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    if (i1 < i0 || i0 < 0 || i1 > n) return -1;
    memset(out3, 0, sizeof(double) * 3 * (size_t)(i1 - i0));
    float boxMin[3], dims[3];
    int bdv[3];
    const int nsrc = orc_domain_bounds(pos4, acc4, i0, i1, NULL, rot, boxMin, dims, bdv);
    if (info) memset(info, 0, sizeof(long long) * 12);
    if (nsrc == 0) return 0;
    if (rot && (bdv[0] != 1 || bdv[1] != 1)) return -4;
    const float rootBoxSize = dims[0];
    const int ndom = bdv[0] * bdv[1] * bdv[2];
    /* domains (setDomainSize :714-752) */
    int* dom = (int*)malloc(sizeof(int) * 3 * (size_t)n);
    int* dcount = (int*)calloc((size_t)ndom, sizeof(int));
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        int nd[3];
        for (int c = 0; c < 3; c++) {
            const float t = (pos4[4 * (size_t)i + c] - boxMin[c]) / rootBoxSize;
            nd[c] = (int)t;
            if (t < 0) nd[c]--;
            if (rootBoxSize == 0) nd[c] = 0;
        }
        for (int c = 0; c < 3; c++)
            if (nd[c] < 0 || nd[c] >= bdv[c]) {
                free(dom);
                free(dcount);
                return -2;
            }
        for (int c = 0; c < 3; c++) dom[3 * (size_t)i + c] = nd[c];
        dcount[(nd[0] * bdv[1] + nd[1]) * bdv[2] + nd[2]]++;
    }
    int maxLevel = 0;
    for (int d = 0; d < ndom; d++) {
        const int l = sp_level_of_count(dcount[d]);
        if (l > maxLevel) maxLevel = l;
    }
    free(dcount);
    if (forced_level != -2) maxLevel = forced_level;
    if (maxLevel < 2) {
        free(dom);
        return -3;
    }
    int bmax = bdv[0];
    if (bdv[1] > bmax) bmax = bdv[1];
    if (bdv[2] > bmax) bmax = bdv[2];
    const int K = sp_ilog2ceil(bmax);
    const int GL = K + maxLevel;  /* leaf level of the enclosing cube */
    const int side = 1 << maxLevel;
    const float boxSizeF = rootBoxSize / (float)side;       /* morton(): float */
    const double boxSizeD = (double)rootBoxSize / side;      /* p2m / l2p: double */

    /* leaf cells (morton :1014-1040) and the sort */
    sp_ki* ki = (sp_ki*)malloc(sizeof(sp_ki) * (size_t)nsrc);
    int ns = 0;
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        int g[3];
        for (int c = 0; c < 3; c++) {
            const int nd = dom[3 * (size_t)i + c];
            const float dmin = boxMin[c] + (float)nd * rootBoxSize;
            int nl = (int)((pos4[4 * (size_t)i + c] - dmin) / boxSizeF);
            if (nl >= side) nl--;
            g[c] = nd * side + nl;
        }
        ki[ns].key = sp_key((uint32_t)g[0], (uint32_t)g[1], (uint32_t)g[2], GL);
        ki[ns].idx = i;
        ns++;
    }
    free(dom);
    qsort(ki, (size_t)ns, sizeof(sp_ki), sp_cmp);

    /* levels */
    sp_level* lv = (sp_level*)calloc((size_t)GL + 1, sizeof(sp_level));
    {
        sp_level* leaf = &lv[GL];
        leaf->key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)ns);
        leaf->first = (int*)malloc(sizeof(int) * (size_t)ns);
        leaf->count = (int*)malloc(sizeof(int) * (size_t)ns);
        int nb = 0;
        for (int a = 0; a < ns; a++) {
            if (a == 0 || ki[a].key != ki[a - 1].key) {
                leaf->key[nb] = ki[a].key;
                leaf->first[nb] = a;
                leaf->count[nb] = 0;
                nb++;
            }
            leaf->count[nb - 1]++;
        }
        leaf->nb = nb;
        for (int l = GL - 1; l >= 2; l--) {
            sp_level* c = &lv[l + 1];
            sp_level* q = &lv[l];
            q->key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)c->nb);
            int m = 0;
            for (int b = 0; b < c->nb; b++) {
                const uint64_t k = c->key[b] >> 3;
                if (m == 0 || q->key[m - 1] != k) q->key[m++] = k;
            }
            q->nb = m;
        }
        for (int l = 2; l <= GL; l++) {
            lv[l].M = (cz*)calloc((size_t)lv[l].nb * nc, sizeof(cz));
            lv[l].L = (cz*)calloc((size_t)lv[l].nb * nc, sizeof(cz));
        }
    }
    long long m2l_pairs = 0, p2p_pairs = 0, nboxes = 0, p2p_ions = 0;
    for (int l = 2; l <= GL; l++) nboxes += lv[l].nb;

    /* P2M (box centre in float, distance as a float difference, p2m :316-343) */
    sp_level* leaf = &lv[GL];
    for (int b = 0; b < leaf->nb; b++) {
        int gx, gy, gz;
        sp_unkey(leaf->key[b], GL, &gx, &gy, &gz);
        const int g[3] = {gx, gy, gz};
        float cen[3];
        for (int c = 0; c < 3; c++) {
            const int nd = g[c] / side, nl = g[c] % side;
            const float dmin = boxMin[c] + (float)nd * rootBoxSize;
            cen[c] = (float)((double)dmin + (nl + 0.5) * boxSizeD);
        }
        cz* M = leaf->M + (size_t)b * nc;
        for (int a = leaf->first[b]; a < leaf->first[b] + leaf->count[b]; a++) {
            const float* q = pos4 + 4 * (size_t)ki[a].idx;
            const float dx = q[0] - cen[0], dy = q[1] - cen[1], dz = q[2] - cen[2];
            sph_p2m_one(p, (double)q[3], dx, dy, dz, boxSizeD, M);
        }
    }
    /* M2M: children in Morton order into each parent.  Rotational, into a
     * domain (level K): each child's four quadrant images, the child placed at
     * octant (x, y) = (1, 1), (0, 1), (0, 0), (1, 0) of quadrant k with its own
     * z (m2m_kernel_rotations, fmm_cpukernel_rotational.cpp:148-215). */
    for (int l = GL - 1; l >= 2; l--) {
        const sp_level* c = &lv[l + 1];
        sp_level* q = &lv[l];
        int pb = 0;
        for (int b = 0; b < c->nb; b++) {
            const uint64_t pk = c->key[b] >> 3;
            while (q->key[pb] != pk) pb++;
            const int o = (int)(c->key[b] & 7u);
            if (rot && l == K) {
                static const int qx[4] = {1, 0, 0, 1}, qy[4] = {1, 1, 0, 0};
                cz Mq[SP_NCOEF];
                for (int k = 0; k < 4; k++) {
                    sp_phase(p, k, c->M + (size_t)b * nc, Mq);
                    sph_m2m(p, qx[k], qy[k], (o >> 2) & 1, Mq, q->M + (size_t)pb * nc);
                }
            } else {
                sph_m2m(p, (o >> 1) & 1, o & 1, (o >> 2) & 1, c->M + (size_t)b * nc, q->M + (size_t)pb * nc);
            }
        }
    }
    /* downward: L2L then M2L at each level 2 .. GL (updateAccelerationNoSym :487-510) */
    for (int l = 2; l <= GL; l++) {
        sp_level* q = &lv[l];
        const int width = 1 << l;
        const double hsz = (double)rootBoxSize * ldexp(1.0, K - l);
        if (l > 2) {
            const sp_level* up = &lv[l - 1];
            int pb = 0;
            for (int b = 0; b < q->nb; b++) {
                const uint64_t pk = q->key[b] >> 3;
                while (up->key[pb] != pk) pb++;
                const int o = (int)(q->key[b] & 7u);
                sph_l2l(p, (o >> 1) & 1, o & 1, (o >> 2) & 1, up->L + (size_t)pb * nc, q->L + (size_t)b * nc);
            }
        }
        for (int b = 0; b < q->nb; b++) {
            int ix, iy, iz;
            sp_unkey(q->key[b], l, &ix, &iy, &iz);
            const int px = ix >> 1, py = iy >> 1, pz = iz >> 1;
            for (int jpx = px - 1; jpx <= px + 1; jpx++)
                for (int jpy = py - 1; jpy <= py + 1; jpy++)
                    for (int jpz = pz - 1; jpz <= pz + 1; jpz++)
                        for (int jx = 2 * jpx; jx <= 2 * jpx + 1; jx++)
                            for (int jy = 2 * jpy; jy <= 2 * jpy + 1; jy++)
                                for (int jz = 2 * jpz; jz <= 2 * jpz + 1; jz++) {
                                    if (!(jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz))
                                        continue;
                                    int sx = jx, sy = jy, k = 0, lim = width;
                                    if (rot && l > K) {  /* quadrant map in the domain's local x, y */
                                        lim = 1 << (l - K);
                                        k = sp_quadrant(jx, jy, lim, &sx, &sy);
                                    }
                                    if (sx < 0 || sy < 0 || jz < 0 || sx >= lim || sy >= lim || jz >= width) continue;
                                    const int sb = sp_find(q, sp_key((uint32_t)sx, (uint32_t)sy, (uint32_t)jz, l));
                                    if (sb < 0) continue;
                                    if (k) {
                                        cz Mq[SP_NCOEF];
                                        sp_phase(p, k, q->M + (size_t)sb * nc, Mq);
                                        sph_m2l(p, ix - jx, iy - jy, iz - jz, hsz, Mq, q->L + (size_t)b * nc);
                                    } else {
                                        sph_m2l(p, ix - jx, iy - jy, iz - jz, hsz, q->M + (size_t)sb * nc,
                                                q->L + (size_t)b * nc);
                                    }
                                    m2l_pairs++;
                                }
        }
    }
    /* L2P and P2P at the leaves */
    const double kc = orc_k_const();
    const int width = 1 << GL;
    for (int b = 0; b < leaf->nb; b++) {
        int ix, iy, iz;
        sp_unkey(leaf->key[b], GL, &ix, &iy, &iz);
        const int g[3] = {ix, iy, iz};
        float cen[3];
        for (int c = 0; c < 3; c++) {
            const int nd = g[c] / side, nl = g[c] % side;
            const float dmin = boxMin[c] + (float)nd * rootBoxSize;
            cen[c] = (float)((double)dmin + (nl + 0.5) * boxSizeD);
        }
        int nb_list[27], nb_quad[27], nnb = 0;
        for (int jx = ix - 1; jx <= ix + 1; jx++)
            for (int jy = iy - 1; jy <= iy + 1; jy++)
                for (int jz = iz - 1; jz <= iz + 1; jz++) {
                    int sx = jx, sy = jy, k = 0, lim = width;
                    if (rot) {  /* the leaf level is below the domains */
                        lim = side;
                        k = sp_quadrant(jx, jy, lim, &sx, &sy);
                    }
                    if (sx < 0 || sy < 0 || jz < 0 || sx >= lim || sy >= lim || jz >= width) continue;
                    const int sb = sp_find(leaf, sp_key((uint32_t)sx, (uint32_t)sy, (uint32_t)jz, GL));
                    if (sb >= 0) {
                        nb_list[nnb] = sb;
                        nb_quad[nnb++] = k;
                    }
                }
        p2p_pairs += nnb;
        for (int t = 0; t < nnb; t++) p2p_ions += (long long)leaf->count[b] * leaf->count[nb_list[t]];
        for (int a = leaf->first[b]; a < leaf->first[b] + leaf->count[b]; a++) {
            const int i = ki[a].idx;
            if (acc4[4 * (size_t)i + 3] != 1.0f) continue;
            const float* pi = pos4 + 4 * (size_t)i;
            const int sp = (int)vel4[4 * (size_t)i + 3];
            const long double mass = sp == 3 ? 0.0L : orc_species_mass_ld(sp);
            const double ac = (double)((long double)(kc * (double)pi[3]) / mass);
            double tot[3] = {0, 0, 0};
            /* P2P: DirectForceFunction::direct per neighbour box, in list order;
             * rotational: direct_rotational with the source rotated by k pi / 2
             * in float (the self term of k = 0 is zero there, skipped here) */
            for (int t = 0; t < nnb; t++) {
                const int sb = nb_list[t];
                const int k = nb_quad[t];
                const float ck = (float)sp_qrf(k + 1), sk = (float)sp_qrf(k);
                double ai[3] = {0, 0, 0};
                for (int e = leaf->first[sb]; e < leaf->first[sb] + leaf->count[sb]; e++) {
                    const int j = ki[e].idx;
                    if (j == i && k == 0) continue;
                    const float* pj = pos4 + 4 * (size_t)j;
                    float xj = pj[0], yj = pj[1];
                    if (k) {
                        xj = ck * pj[0] - sk * pj[1];
                        yj = sk * pj[0] + ck * pj[1];
                    }
                    const double d0 = (double)(pi[0] - xj);
                    const double d1 = (double)(pi[1] - yj);
                    const double d2 = (double)(pi[2] - pj[2]);
                    const double invD = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2 + SP_SOFTENING);
                    const double invD3 = (double)pj[3] * invD * invD * invD;
                    ai[0] += d0 * invD3;
                    ai[1] += d1 * invD3;
                    ai[2] += d2 * invD3;
                }
                for (int c = 0; c < 3; c++) tot[c] += ac * ai[c];
            }
            double f[3];
            const float dx = pi[0] - cen[0], dy = pi[1] - cen[1], dz = pi[2] - cen[2];
            sph_l2p_one(p, leaf->L + (size_t)b * nc, dx, dy, dz, boxSizeD, f);
            for (int c = 0; c < 3; c++) out3[3 * (size_t)(i - i0) + c] = tot[c] - ac * f[c];
        }
    }
    if (info) {
        info[0] = maxLevel;
        info[1] = K;
        info[2] = bdv[0];
        info[3] = bdv[1];
        info[4] = bdv[2];
        info[5] = leaf->nb;
        info[6] = m2l_pairs;
        info[7] = p2p_pairs;
        info[8] = ns;
        info[9] = nboxes;
        info[10] = p2p_ions;
    }
    for (int l = 2; l <= GL; l++) {
        free(lv[l].key);
        free(lv[l].M);
        free(lv[l].L);
    }
    free(leaf->first);
    free(leaf->count);
    free(lv);
    free(ki);
    return 0;
}

/* ------------------------------------------------------ translational FMM -- */

/* Base-27 cell index of the ternary tree (morton3, multipole_method_functions.cpp:49-65):
 * level digit 9 z + 3 x + y, finest level least significant. */
static uint64_t sp_tkey(int x, int y, int z, int levels) {
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
/* unmorton3 (multipole_method_functions.cpp:68-86). */
static void sp_tunkey(uint64_t k, int* x, int* y, int* z) {
    int a = 0, b = 0, c = 0, w = 1;
    while (k) {
        const int d = (int)(k % 27);
        k /= 27;
        b += (d % 3) * w;
        a += ((d / 3) % 3) * w;
        c += (d / 9) * w;
        w *= 3;
    }
    *x = a;
    *y = b;
    *z = c;
}
static int sp_wrap(int v, int w, int* shift) {
    *shift = 0;
    while (v < 0) { v += w; (*shift)--; }
    while (v >= w) { v -= w; (*shift)++; }
    return v;
}

/*
 * The reference's translational FMM (updateAccelerationTranslational,
 * multipole_method_force_calculation_translational.cpp:9-331) for a cubic
 * periodic cell of side `cell` on all three axes (symmetryAxes (1, 1, 1),
 * numSymmetryLevels >= 1): one domain, boxMin = -cell / 2 on every axis
 * (computeDomainBoundsTranslations, domain_bounds_functions.cpp:637-661),
 * a ternary tree of levels 1 .. maxLevel (3^l cells per axis, float leaf
 * placement of mortonTranslations :669-715, level from setOptimumLevel-
 * Translations :521-567), lists that wrap periodically (getInteraction-
 * ListTranslations :1217-1334 P2P, :1725-1834 M2L_Subcell) with the image
 * shift in the P2P list, and the operators of fmm_cpukernel_translational.cpp:
 * P2M :344-446, M2M :455-558 (ternary shift, 3^-j), M2L :567-748 (offsets up
 * to 5 boxes, gcd-reduced rotation), L2L :756-835 (3^(-n-1)), L2P :843-973.
 * Every target then sees the sources of the 27 cells at shifts {-1, 0, 1}^3.
 * near != 0: the P2P adds direct_translational (direct_force_method_
 * translational.cpp:320-367) over each listed leaf pair, the target shifted
 * by -shift * cell in float.  near == 0: the reference as it runs, where
 * p2p_kernel_translations (fmm_cpukernel_translational.cpp:305-336) passes
 * the thread's box range in place of the leaf pair's ion bounds, so the
 * source loop (j from numBoxIndex up to 0) is empty and P2P adds nothing.
 * out3, info and the ranges as orc_fmm_sph (info[1] = 0, [2..4] = 1).
 * Returns 0; -1 bad arguments; -2 a source outside the cell (the
 * reference's outlier passes); -3 maxLevel < 1 (its supercell-only path)
 * or > 6 (its int box index overflows at 27^7).
 */
int orc_fmm_sph_trans(const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1, int p,
                      int forced_level, float cell, int near, double* out3, long long* info) {
    if (p < 1 || p > SP_PMAX || i1 < i0 || i0 < 0 || i1 > n || !(cell > 0)) return -1;
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    memset(out3, 0, sizeof(double) * 3 * (size_t)(i1 - i0));
    if (info) memset(info, 0, sizeof(long long) * 12);
    const float rootBoxSize = cell;
    const float bmin = -rootBoxSize / 2;
    int nsrc = 0;
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        for (int c = 0; c < 3; c++) {
            const float t = (pos4[4 * (size_t)i + c] - bmin) / rootBoxSize;
            int v = (int)t;
            if (t < 0) v--;
            if (v != 0) return -2;
        }
        nsrc++;
    }
    if (nsrc == 0) return 0;
    int maxLevel = sp_level_of_count(nsrc);
    if (forced_level != -2) maxLevel = forced_level;
    if (maxLevel < 1 || maxLevel > 6) return -3;
    const int GL = maxLevel;
    int side = 1;
    for (int l = 0; l < GL; l++) side *= 3;
    const float numTotalBoxes = (float)pow(3.0, GL);
    const float boxSizeF = rootBoxSize / numTotalBoxes;          /* mortonTranslations: float */
    const double boxSizeD = rootBoxSize / pow(3.0, GL);          /* p2m / l2p: double */

    sp_ki* ki = (sp_ki*)malloc(sizeof(sp_ki) * (size_t)nsrc);
    int ns = 0;
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        int g[3];
        for (int c = 0; c < 3; c++) {
            int v = (int)((pos4[4 * (size_t)i + c] - bmin) / boxSizeF);
            if (v >= numTotalBoxes) v--;
            g[c] = v;
        }
        ki[ns].key = sp_tkey(g[0], g[1], g[2], GL);
        ki[ns].idx = i;
        ns++;
    }
    qsort(ki, (size_t)ns, sizeof(sp_ki), sp_cmp);

    sp_level* lv = (sp_level*)calloc((size_t)GL + 1, sizeof(sp_level));
    sp_level* leaf = &lv[GL];
    leaf->key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)ns);
    leaf->first = (int*)malloc(sizeof(int) * (size_t)ns);
    leaf->count = (int*)malloc(sizeof(int) * (size_t)ns);
    {
        int nb = 0;
        for (int a = 0; a < ns; a++) {
            if (a == 0 || ki[a].key != ki[a - 1].key) {
                leaf->key[nb] = ki[a].key;
                leaf->first[nb] = a;
                leaf->count[nb] = 0;
                nb++;
            }
            leaf->count[nb - 1]++;
        }
        leaf->nb = nb;
    }
    for (int l = GL - 1; l >= 1; l--) {
        sp_level* c = &lv[l + 1];
        sp_level* q = &lv[l];
        q->key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)c->nb);
        int m = 0;
        for (int b = 0; b < c->nb; b++) {
            const uint64_t k = c->key[b] / 27;
            if (m == 0 || q->key[m - 1] != k) q->key[m++] = k;
        }
        q->nb = m;
    }
    long long m2l_pairs = 0, p2p_pairs = 0, nboxes = 0, p2p_ions = 0;
    for (int l = 1; l <= GL; l++) {
        lv[l].M = (cz*)calloc((size_t)lv[l].nb * nc, sizeof(cz));
        lv[l].L = (cz*)calloc((size_t)lv[l].nb * nc, sizeof(cz));
        nboxes += lv[l].nb;
    }
    /* P2M: the leaf centre as p2m_kernel_translations forms it (:387-392) */
    for (int b = 0; b < leaf->nb; b++) {
        int g[3];
        sp_tunkey(leaf->key[b], &g[0], &g[1], &g[2]);
        float cen[3];
        for (int c = 0; c < 3; c++) cen[c] = (float)((double)bmin + (g[c] + 0.5) * boxSizeD);
        cz* M = leaf->M + (size_t)b * nc;
        for (int a = leaf->first[b]; a < leaf->first[b] + leaf->count[b]; a++) {
            const float* q = pos4 + 4 * (size_t)ki[a].idx;
            const float dx = q[0] - cen[0], dy = q[1] - cen[1], dz = q[2] - cen[2];
            sph_p2m_one(p, (double)q[3], dx, dy, dz, boxSizeD, M);
        }
    }
    /* M2M down to level 1 (the loop of :180-210 stops at max(0, 2 - numLevelSC) = 1) */
    for (int l = GL - 1; l >= 1; l--) {
        const sp_level* c = &lv[l + 1];
        sp_level* q = &lv[l];
        int pb = 0;
        for (int b = 0; b < c->nb; b++) {
            const uint64_t pk = c->key[b] / 27;
            while (q->key[pb] != pk) pb++;
            int x, y, z;
            sp_tunkey(c->key[b] % 27, &x, &y, &z);
            const double rho = sqrt((double)(abs(x - 1) + abs(y - 1) + abs(z - 1)));
            sph_m2m_g(p, 6 - x, 6 - y, 6 - z, rho, 3.0, c->M + (size_t)b * nc, q->M + (size_t)pb * nc);
        }
    }
    /* downward: M2L at level 1, then L2L and M2L at 2 .. maxLevel (:242-313) */
    for (int l = 1; l <= GL; l++) {
        sp_level* q = &lv[l];
        int width = 1;
        for (int i = 0; i < l; i++) width *= 3;
        const double hsz = rootBoxSize / pow(3.0, l);
        if (l > 1) {
            const sp_level* up = &lv[l - 1];
            int pb = 0;
            for (int b = 0; b < q->nb; b++) {
                const uint64_t pk = q->key[b] / 27;
                while (up->key[pb] != pk) pb++;
                int x, y, z;
                sp_tunkey(q->key[b] % 27, &x, &y, &z);
                const double rho = sqrt((double)(abs(x - 1) + abs(y - 1) + abs(z - 1)));
                sph_l2l_g(p, x + 4, y + 4, z + 4, rho, 3.0, up->L + (size_t)pb * nc, q->L + (size_t)b * nc);
            }
        }
        for (int b = 0; b < q->nb; b++) {
            int ix, iy, iz;
            sp_tunkey(q->key[b], &ix, &iy, &iz);
            const int px = ix / 3, py = iy / 3, pz = iz / 3;
            for (int jpx = px - 1; jpx <= px + 1; jpx++)
                for (int jpy = py - 1; jpy <= py + 1; jpy++)
                    for (int jpz = pz - 1; jpz <= pz + 1; jpz++)
                        for (int jx = 3 * jpx; jx <= 3 * jpx + 2; jx++)
                            for (int jy = 3 * jpy; jy <= 3 * jpy + 2; jy++)
                                for (int jz = 3 * jpz; jz <= 3 * jpz + 2; jz++) {
                                    if (!(jx < ix - 1 || ix + 1 < jx || jy < iy - 1 || iy + 1 < jy || jz < iz - 1 || iz + 1 < jz))
                                        continue;
                                    int s0, s1, s2;
                                    const int sx = sp_wrap(jx, width, &s0), sy = sp_wrap(jy, width, &s1),
                                              sz = sp_wrap(jz, width, &s2);
                                    const int sb = sp_find(q, sp_tkey(sx, sy, sz, l));
                                    if (sb < 0) continue;
                                    sph_m2l(p, ix - jx, iy - jy, iz - jz, hsz, q->M + (size_t)sb * nc,
                                            q->L + (size_t)b * nc);
                                    m2l_pairs++;
                                }
        }
    }
    /* L2P and P2P at the leaves */
    const double kc = orc_k_const();
    for (int b = 0; b < leaf->nb; b++) {
        int ix, iy, iz;
        sp_tunkey(leaf->key[b], &ix, &iy, &iz);
        const int g[3] = {ix, iy, iz};
        float cen[3];
        for (int c = 0; c < 3; c++) cen[c] = (float)((double)bmin + (g[c] + 0.5) * boxSizeD);
        int nb_list[27], nb_shift[27][3], nnb = 0;
        for (int jx = ix - 1; jx <= ix + 1; jx++)
            for (int jy = iy - 1; jy <= iy + 1; jy++)
                for (int jz = iz - 1; jz <= iz + 1; jz++) {
                    int s0, s1, s2;
                    const int sx = sp_wrap(jx, side, &s0), sy = sp_wrap(jy, side, &s1), sz = sp_wrap(jz, side, &s2);
                    const int sb = sp_find(leaf, sp_tkey(sx, sy, sz, GL));
                    if (sb < 0) continue;
                    nb_list[nnb] = sb;
                    nb_shift[nnb][0] = s0;
                    nb_shift[nnb][1] = s1;
                    nb_shift[nnb][2] = s2;
                    nnb++;
                }
        p2p_pairs += nnb;
        for (int t = 0; t < nnb; t++) p2p_ions += (long long)leaf->count[b] * leaf->count[nb_list[t]];
        for (int a = leaf->first[b]; a < leaf->first[b] + leaf->count[b]; a++) {
            const int i = ki[a].idx;
            if (acc4[4 * (size_t)i + 3] != 1.0f) continue;
            const float* pi = pos4 + 4 * (size_t)i;
            const int sp = (int)vel4[4 * (size_t)i + 3];
            const long double mass = sp == 3 ? 0.0L : orc_species_mass_ld(sp);
            const double ac = (double)((long double)(kc * (double)pi[3]) / mass);
            double tot[3] = {0, 0, 0};
            for (int t = 0; near && t < nnb; t++) {
                const int sb = nb_list[t];
                const float off[3] = {(float)nb_shift[t][0] * cell, (float)nb_shift[t][1] * cell,
                                      (float)nb_shift[t][2] * cell};
                const float xi = pi[0] - off[0], yi = pi[1] - off[1], zi = pi[2] - off[2];
                double ai[3] = {0, 0, 0};
                for (int e = leaf->first[sb]; e < leaf->first[sb] + leaf->count[sb]; e++) {
                    const float* pj = pos4 + 4 * (size_t)ki[e].idx;
                    const double d0 = (double)(xi - pj[0]);
                    const double d1 = (double)(yi - pj[1]);
                    const double d2 = (double)(zi - pj[2]);
                    const double invD = 1.0 / sqrt(d0 * d0 + d1 * d1 + d2 * d2 + SP_SOFTENING);
                    const double invD3 = (double)pj[3] * invD * invD * invD;
                    ai[0] += d0 * invD3;
                    ai[1] += d1 * invD3;
                    ai[2] += d2 * invD3;
                }
                for (int c = 0; c < 3; c++) tot[c] += ac * ai[c];
            }
            double f[3];
            const float dx = pi[0] - cen[0], dy = pi[1] - cen[1], dz = pi[2] - cen[2];
            sph_l2p_one(p, leaf->L + (size_t)b * nc, dx, dy, dz, boxSizeD, f);
            for (int c = 0; c < 3; c++) out3[3 * (size_t)(i - i0) + c] = tot[c] - ac * f[c];
        }
    }
    if (info) {
        info[0] = maxLevel;
        info[2] = info[3] = info[4] = 1;
        info[5] = leaf->nb;
        info[6] = m2l_pairs;
        info[7] = p2p_pairs;
        info[8] = ns;
        info[9] = nboxes;
        info[10] = p2p_ions;
    }
    for (int l = 1; l <= GL; l++) {
        free(lv[l].key);
        free(lv[l].M);
        free(lv[l].L);
    }
    free(leaf->first);
    free(leaf->count);
    free(lv);
    free(ki);
    return 0;
}

/* ----------------------------------------------------------- unit access -- */
/* Expansions cross these entry points as interleaved (re, im) doubles. */

static void to_cz(const double* v, cz* c, int nc) {
    for (int i = 0; i < nc; i++) c[i] = cz_make(v[2 * i], v[2 * i + 1]);
}
static void from_cz(const cz* c, double* v, int nc) {
    for (int i = 0; i < nc; i++) {
        v[2 * i] = c[i].re;
        v[2 * i + 1] = c[i].im;
    }
}

/* P2M of m particles (x, y, z, q as doubles relative to the box centre) into M. */
void orc_sph_p2m(int p, int m, const double* xyzq, double box_size, double* M) {
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    cz c[SP_NCOEF];
    to_cz(M, c, nc);
    for (int i = 0; i < m; i++) sph_p2m_one(p, xyzq[4 * i + 3], xyzq[4 * i], xyzq[4 * i + 1], xyzq[4 * i + 2], box_size, c);
    from_cz(c, M, nc);
}
/* Child expansion of octant o (reference Morton digit) added into the parent's. */
void orc_sph_m2m(int p, int o, const double* Mc, double* Mp) {
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    cz a[SP_NCOEF], b[SP_NCOEF];
    to_cz(Mc, a, nc);
    to_cz(Mp, b, nc);
    sph_m2m(p, (o >> 1) & 1, o & 1, (o >> 2) & 1, a, b);
    from_cz(b, Mp, nc);
}
/* Source box at offset (dx, dy, dz) = target - source, boxes of size h. */
void orc_sph_m2l(int p, int dx, int dy, int dz, double h, const double* M, double* L) {
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    cz a[SP_NCOEF], b[SP_NCOEF];
    to_cz(M, a, nc);
    to_cz(L, b, nc);
    sph_m2l(p, dx, dy, dz, h, a, b);
    from_cz(b, L, nc);
}
void orc_sph_l2l(int p, int o, const double* Lp, double* Lc) {
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    cz a[SP_NCOEF], b[SP_NCOEF];
    to_cz(Lp, a, nc);
    sph_l2l(p, (o >> 1) & 1, o & 1, (o >> 2) & 1, a, b);
    from_cz(b, Lc, nc);
}
void orc_sph_l2p(int p, const double* L, double dx, double dy, double dz, double h, double* f3) {
    sph_precalc();
    const int nc = p * (p + 1) / 2;
    cz a[SP_NCOEF];
    to_cz(L, a, nc);
    sph_l2p_one(p, a, dx, dy, dz, h, f3);
}
/* The rotation of direction (dx, dy, dz) in [-5, 5]^3 with gcd 1 (forward)
 * or of its x-mirror (backward), as the M2L applies them. */
int orc_sph_rotate(int p, int dx, int dy, int dz, int backward, const double* C, double* out) {
    sph_precalc();
    if (gcd3(dx, dy, dz) != 1) return -1;
    int fwd, bwd;
    rotation_index(dx + 5, dy + 5, dz + 5, &fwd, &bwd);
    const int nc = p * (p + 1) / 2;
    cz a[SP_NCOEF], b[SP_NCOEF];
    to_cz(C, a, nc);
    sph_rotate(p, a, b, backward ? bwd : fwd);
    from_cz(b, out, nc);
    return 0;
}
/* Table values for the tests: which 0 factorial, 1 anm, 2 Anm; 3 Ynm (re). */
double orc_sph_table(int which, int index) {
    sph_precalc();
    switch (which) {
    case 0: return g_factorial[index];
    case 1: return g_anm[index];
    case 2: return g_Anm[index];
    case 3: return g_Ynm[index].re;
    }
    return 0;
}
