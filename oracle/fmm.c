/*
 * oracle/fmm.c -- CPU FP64 restatement of the engine's fast multipole method.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/direct.c).
 *
 * The reference FMM (multipole_method_force_calculation.cpp, fmm_cpukernel.cpp)
 * does not compile as shipped and uses rotation-based spherical-harmonic
 * translations of the plain 1/r kernel with a softened centre distance in M2L
 * (fmm_cpukernel.cpp:832).  The engine instead expands the softened kernel
 * G(r) = (|r|^2 + eps^2)^(-1/2) itself in Cartesian Taylor series, so the far
 * field converges to the softened direct sum (direct_force_method.cpp:170-183).
 * This file is the double-precision statement of exactly that algorithm, used
 * to check the GPU passes one by one; its accuracy against orc_direct_f64 is
 * checked separately.
 *
 * Tree: cube root from the bounding box of the active particles; uniform leaf
 * level L; box at level l = Morton prefix of the particle's level-L cell.
 * Octree Morton bit order x -> bit 3i, y -> 3i+1, z -> 3i+2.
 * Near field: boxes whose integer cell coordinates differ by at most `ring` in
 * every axis (P2P at the leaves).  Far field: at each level l >= 1, children of
 * the parent's near boxes that are not near themselves (M2L).
 *
 * Expansions, with h = box half size and hats for lengths in units of h:
 *   P2M  M_a = sum_j q_j v^a,            v = (y_j - c)/h
 *   M2M  M'_a = sum_{k<=a} C(a,k) d^(a-k) 2^-|k| M_k,  d = (c_child - c_parent)/h_parent
 *   M2L  L_b += sum_{|a|<=p-|b|} C(a+b,a) (-1)^|a| T_{a+b}(R/h, eps/h) M_a
 *   L2L  L'_k = 1/2 sum_{b>=k} C(b,k) 2^-|k| e^(b-k) L_b,  e = (c_child - c_parent)/h_parent
 *   L2P  phi(x) = (1/h) sum_b L_b u^b,   u = (x - c)/h
 * T_g = d^g G / g! obeys, with s = |R|^2 + eps^2 (Duan and Krasny, J. Comput.
 * Chem. 22 (2001) 184):
 *   |g| s T_g + (2|g|-1) sum_i R_i T_{g-e_i} + (|g|-1) sum_i T_{g-2e_i} = 0.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Threads of the box loops below (0: OpenMP's default).  Every loop runs
 * over target boxes that own their outputs, so the result does not depend on
 * the thread count. */
static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }

/* L2P + P2P on every k-th leaf box only (1: all), for bounded timing samples;
 * the far-field passes always run over every box. */
static int g_leaf_stride = 1;
void orc_fmm_set_leaf_stride(int k) { g_leaf_stride = k > 1 ? k : 1; }
int orc_get_threads(void) { return g_threads > 0 ? g_threads : omp_get_max_threads(); }

void orc_direct_f64(int n, const double* x, const double* y, const double* z, const double* q,
                    const int* status, const int* species, int i0, int i1, int sym, int nsec,
                    int nimg, const int* lattice, const double* domain, double* out3);
double orc_k_const(void);
long double orc_species_mass_ld(int s);

#define MAXP 20

typedef struct {
    int p, nc;
    int (*mi)[3];                          /* multi-indices by degree */
    int idx[MAXP + 1][MAXP + 1][MAXP + 1]; /* (a,b,c) -> index */
} Basis;

static void basis_init(Basis* B, int p) {
    B->p = p;
    B->nc = (p + 1) * (p + 2) * (p + 3) / 6;
    B->mi = malloc(sizeof(int[3]) * B->nc);
    int k = 0;
    for (int n = 0; n <= p; n++)
        for (int a = n; a >= 0; a--)
            for (int b = n - a; b >= 0; b--) {
                int c = n - a - b;
                B->mi[k][0] = a; B->mi[k][1] = b; B->mi[k][2] = c;
                B->idx[a][b][c] = k;
                k++;
            }
}

int orc_fmm_ncoef(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

static double binom(int n, int k) {
    double r = 1.0;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}

/* T_g for |g| <= B->p at R (lengths already scaled), eps2 = eps^2 scaled. */
static void taylor_b(const Basis* Bp, const double R[3], double eps2, double* T) {
    const Basis* B = Bp;
    double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps2;
    T[0] = 1.0 / sqrt(s);
    for (int k = 1; k < B->nc; k++) {
        const int* g = B->mi[k];
        int n = g[0] + g[1] + g[2];
        double acc = 0.0;
        for (int i = 0; i < 3; i++) {
            if (g[i] >= 1) {
                int h[3] = {g[0], g[1], g[2]};
                h[i] -= 1;
                acc += (2 * n - 1) * R[i] * T[B->idx[h[0]][h[1]][h[2]]];
            }
            if (g[i] >= 2) {
                int h[3] = {g[0], g[1], g[2]};
                h[i] -= 2;
                acc += (n - 1) * T[B->idx[h[0]][h[1]][h[2]]];
            }
        }
        T[k] = -acc / (n * s);
    }
}

void orc_fmm_taylor(int p, const double R[3], double eps2, double* T) {
    Basis B;
    basis_init(&B, p);
    double s = R[0] * R[0] + R[1] * R[1] + R[2] * R[2] + eps2;
    T[0] = 1.0 / sqrt(s);
    for (int k = 1; k < B.nc; k++) {
        int* g = B.mi[k];
        int n = g[0] + g[1] + g[2];
        double acc = 0.0;
        for (int i = 0; i < 3; i++) {
            if (g[i] >= 1) {
                int h[3] = {g[0], g[1], g[2]};
                h[i] -= 1;
                acc += (2 * n - 1) * R[i] * T[B.idx[h[0]][h[1]][h[2]]];
            }
            if (g[i] >= 2) {
                int h[3] = {g[0], g[1], g[2]};
                h[i] -= 2;
                acc += (n - 1) * T[B.idx[h[0]][h[1]][h[2]]];
            }
        }
        T[k] = -acc / (n * s);
    }
    free(B.mi);
}

static double ipow(double x, int e) {
    double r = 1.0;
    for (int i = 0; i < e; i++) r *= x;
    return r;
}

/* ------------------------------------------------------------------ tree --- */

typedef struct {
    int L, ring, p;
    double x0[3], side;          /* root cube */
    int nlev;                    /* levels 0..L */
    int* nbox;                   /* boxes per level */
    uint64_t** key;              /* sorted box keys per level */
    int** first;                 /* first particle (leaf level) or first child */
    int** count;                 /* particles (leaf) or children */
    double** M;                  /* multipoles [box][nc] */
    double** Lc;                 /* locals */
} Tree;

typedef struct { uint64_t k; int i; } KI;  /* leaf key and particle of the sorted sources */

static uint64_t spread21(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
static uint64_t compact21(uint64_t v) {
    v &= 0x1249249249249249ULL;
    v = (v ^ (v >> 2)) & 0x10c30c30c30c30c3ULL;
    v = (v ^ (v >> 4)) & 0x100f00f00f00f00fULL;
    v = (v ^ (v >> 8)) & 0x1f0000ff0000ffULL;
    v = (v ^ (v >> 16)) & 0x1f00000000ffffULL;
    v = (v ^ (v >> 32)) & 0x1fffff;
    return v;
}

/* Root cube of the active particles: side = largest extent times (1 + 2^-20),
 * centred on the bounding box.  Cell of a coordinate at level L:
 * floor((x - x0) * (2^L / side)) clamped to [0, 2^L - 1]. */
void orc_fmm_root(int n, const double* x, const double* y, const double* z, const int* status,
                  double out[4]) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    const double* P[3] = {x, y, z};
    for (int i = 0; i < n; i++) {
        if (!(status[i] == 1 || status[i] == 2)) continue;
        for (int c = 0; c < 3; c++) {
            mn[c] = fmin(mn[c], P[c][i]);
            mx[c] = fmax(mx[c], P[c][i]);
        }
    }
    double side = fmax(fmax(mx[0] - mn[0], mx[1] - mn[1]), mx[2] - mn[2]);
    if (!(side > 0)) side = 1e-9;
    side = side * (1.0 + 0x1p-20);
    for (int c = 0; c < 3; c++) out[c] = 0.5 * (mn[c] + mx[c]) - 0.5 * side;
    out[3] = side;
}

uint64_t orc_fmm_key(double x, double y, double z, const double root[4], int L) {
    double sc = ldexp(1.0, L) / root[3];
    double v[3] = {x, y, z};
    uint64_t k = 0;
    int64_t top = ((int64_t)1 << L) - 1;
    for (int c = 0; c < 3; c++) {
        double f = floor((v[c] - root[c]) * sc);
        int64_t iv = (int64_t)f;
        if (f < 0) iv = 0;
        if (iv > top) iv = top;
        k |= spread21((uint64_t)iv) << c;
    }
    return k;
}

static int find_box(const Tree* t, int l, uint64_t key) {
    int lo = 0, hi = t->nbox[l] - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (t->key[l][mid] == key) return mid;
        if (t->key[l][mid] < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}

static void box_center(const Tree* t, int l, uint64_t key, double c[3], double* h) {
    double bs = t->side / ldexp(1.0, l);
    for (int d = 0; d < 3; d++) c[d] = t->x0[d] + (compact21(key >> d) + 0.5) * bs;
    *h = 0.5 * bs;
}

/* Field of box b of level l at X from its multipole expansion (M2P,
 * fmm_cpukernel.cpp:1319-1533):  with u = (X - c)/h,
 *   acc_k += -(1/h^2) sum_{|a|<=p} (-1)^|a| M_a (a_k + 1) T_{a+e_k}(u),
 * B the basis of order p, B1 of order p + 1, T scratch of B1->nc. */
static void m2p_box(const Tree* t, const Basis* B, const Basis* B1, int l, int b, const double X[3], double eps2,
                    double* T, double acc[3]) {
    double c[3], h;
    box_center(t, l, t->key[l][b], c, &h);
    double u[3] = {(X[0] - c[0]) / h, (X[1] - c[1]) / h, (X[2] - c[2]) / h};
    taylor_b(B1, u, eps2 / (h * h), T);
    const double* M = t->M[l] + (size_t)b * B->nc;
    double g[3] = {0, 0, 0};
    for (int a = 0; a < B->nc; a++) {
        const int* A = B->mi[a];
        const double sg = ((A[0] + A[1] + A[2]) & 1) ? -M[a] : M[a];
        g[0] += sg * (A[0] + 1) * T[B1->idx[A[0] + 1][A[1]][A[2]]];
        g[1] += sg * (A[1] + 1) * T[B1->idx[A[0]][A[1] + 1][A[2]]];
        g[2] += sg * (A[2] + 1) * T[B1->idx[A[0]][A[1]][A[2] + 1]];
    }
    for (int k = 0; k < 3; k++) acc[k] -= g[k] / (h * h);
}

static int64_t floor_half(int64_t v) { return v >= 0 ? v / 2 : -((-v + 1) / 2); }

/* The tree's field at the point X: at each level l = 0..L the boxes whose
 * parent is near X's cell one level up (the root at level 0) and that are not
 * near X's own cell floor((X - x0) / w_l) (not clamped: X may lie outside the
 * root cube) by M2P; with near_p2p, the ions of the leaf boxes near X's leaf
 * cell directly.  Adds sum q d / r^3 into acc; returns the M2P count. */
static double m2p_point(const Tree* t, const Basis* B, const Basis* B1, const KI* ki, const double* x,
                        const double* y, const double* z, const double* q, const double X[3], double eps2,
                        int near_p2p, double* T, double acc[3]) {
    if (t->nbox[0] == 0) return 0;
    double nm2p = 0;
    const int ring = t->ring;
    for (int l = 0; l <= t->L; l++) {
        const double w = t->side / ldexp(1.0, l);
        int64_t cc[3];
        for (int d = 0; d < 3; d++) cc[d] = (int64_t)floor((X[d] - t->x0[d]) / w);
        if (l == 0) {
            int near = llabs(cc[0]) <= ring && llabs(cc[1]) <= ring && llabs(cc[2]) <= ring;
            if (!near) {
                m2p_box(t, B, B1, 0, 0, X, eps2, T, acc);
                return nm2p + 1;
            }
            if (t->L == 0 && near_p2p) {
                for (int s = t->first[0][0]; s < t->first[0][0] + t->count[0][0]; s++) {
                    int j = ki[s].i;
                    double dx = X[0] - x[j], dy = X[1] - y[j], dz = X[2] - z[j];
                    double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps2);
                    double wq = q[j] * inv * inv * inv;
                    acc[0] += dx * wq; acc[1] += dy * wq; acc[2] += dz * wq;
                }
            }
            continue;
        }
        const int64_t qc[3] = {floor_half(cc[0]), floor_half(cc[1]), floor_half(cc[2])};
        const int64_t ptop = ((int64_t)1 << (l - 1)) - 1;
        for (int64_t dz = -ring; dz <= ring; dz++)
            for (int64_t dy = -ring; dy <= ring; dy++)
                for (int64_t dx = -ring; dx <= ring; dx++) {
                    int64_t pc[3] = {qc[0] + dx, qc[1] + dy, qc[2] + dz};
                    if (pc[0] < 0 || pc[1] < 0 || pc[2] < 0 || pc[0] > ptop || pc[1] > ptop || pc[2] > ptop) continue;
                    for (int oc = 0; oc < 8; oc++) {
                        int64_t bc[3] = {2 * pc[0] + (oc & 1), 2 * pc[1] + ((oc >> 1) & 1), 2 * pc[2] + ((oc >> 2) & 1)};
                        int near = llabs(bc[0] - cc[0]) <= ring && llabs(bc[1] - cc[1]) <= ring &&
                                   llabs(bc[2] - cc[2]) <= ring;
                        if (near && (l < t->L || !near_p2p)) continue;
                        uint64_t key = spread21(bc[0]) | spread21(bc[1]) << 1 | spread21(bc[2]) << 2;
                        int b = find_box(t, l, key);
                        if (b < 0) continue;
                        if (!near) {
                            m2p_box(t, B, B1, l, b, X, eps2, T, acc);
                            nm2p += 1;
                        } else {
                            for (int s = t->first[l][b]; s < t->first[l][b] + t->count[l][b]; s++) {
                                int j = ki[s].i;
                                double ex = X[0] - x[j], ey = X[1] - y[j], ez = X[2] - z[j];
                                double inv = 1.0 / sqrt(ex * ex + ey * ey + ez * ez + eps2);
                                double wq = q[j] * inv * inv * inv;
                                acc[0] += ex * wq; acc[1] += ey * wq; acc[2] += ez * wq;
                            }
                        }
                    }
                }
    }
    return nm2p;
}

/* Optional M2L order per Chebyshev shell of the box offset (0 = use p). */
static int g_pshell[8];
void orc_fmm_set_shell_orders(const int* ps, int n) {
    for (int k = 0; k < 8; k++) g_pshell[k] = (ps && k < n) ? ps[k] : 0;
}

/* FmmType (structures.h:283-288): 0 treecode, 1 treecode with outliers,
 * 2 fmm (default), 3 fmm with outliers. */
static int g_fmm_type = 2;
void orc_fmm_set_type(int type) { g_fmm_type = type; }

/* DomainBoundsFunction::isOutlier (domain_bounds_functions.cpp:16-93) as a
 * defined statistic: per axis the median of the valid ions (status 1 or 2; the
 * mean of the two middle values for an even count, findTrueMedian,
 * outlier_detection.h:118-133) and their median absolute deviation times
 * outlierFactor (1.4826 as a float); flag[i] = 1 for a valid ion whose
 * deviation is at least 3 such units on some axis.  Returns the count. */
static int cmp_dbl(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}
static double median_sorted(const double* v, int m) { return m > 0 ? 0.5 * (v[(m - 1) / 2] + v[m / 2]) : 0.0; }

int orc_fmm_outliers(int n, const double* x, const double* y, const double* z, const int* status, int* flag) {
    const double* P[3] = {x, y, z};
    double* v = malloc(sizeof(double) * (n > 0 ? n : 1));
    double med[3], mad[3];
    for (int c = 0; c < 3; c++) {
        int m = 0;
        for (int i = 0; i < n; i++)
            if (status[i] == 1 || status[i] == 2) v[m++] = P[c][i];
        qsort(v, m, sizeof(double), cmp_dbl);
        med[c] = median_sorted(v, m);
        m = 0;
        for (int i = 0; i < n; i++)
            if (status[i] == 1 || status[i] == 2) v[m++] = fabs(P[c][i] - med[c]);
        qsort(v, m, sizeof(double), cmp_dbl);
        mad[c] = (double)1.482602218505602f * median_sorted(v, m);
    }
    int nout = 0;
    for (int i = 0; i < n; i++) {
        int out = 0;
        if (status[i] == 1 || status[i] == 2)
            for (int c = 0; c < 3; c++) out |= fabs(P[c][i] - med[c]) / mad[c] >= 3.0;
        flag[i] = out;
        nout += out;
    }
    free(v);
    return nout;
}

static int cmp_ki(const void* a, const void* b) {
    const KI* x = a;
    const KI* y = b;
    if (x->k != y->k) return x->k < y->k ? -1 : 1;
    return x->i - y->i;
}

/*
 * FMM accelerations for the active particles (status 1 targets, 1/2 sources).
 * Outputs out3 (double, zero for non-targets).  p: expansion order, L: leaf
 * level, ring: near-field width.  stats (may be 0) receives
 * {boxes at leaf level, P2P pairs, M2L interactions, seconds of the tree and
 * far-field passes (sort to L2L), seconds of L2P + P2P, leaf boxes evaluated,
 * M2P evaluations}.
 * orc_fmm_set_type selects FmmType (structures.h:283-288): the treecode types
 * skip M2L and L2L and give every target the M2P field of its far boxes at
 * every level (m2p_point); the outlier types hold the orc_fmm_outliers ions
 * out of the tree, give them the tree's field by M2P plus their near leaf
 * cells, and give every target the outliers' field directly.
 */
void orc_fmm_f64(int n, const double* x, const double* y, const double* z, const double* q,
                 const int* status_in, const int* species, int p, int L, int ring, double* out3,
                 double* stats) {
    const double t_start = omp_get_wtime();
    const int treecode = g_fmm_type == 0 || g_fmm_type == 1;
    const int outliers = g_fmm_type == 1 || g_fmm_type == 3;
    /* outlier modes: the tree holds the other ions (outliers get status 0 in it) */
    int* oflag = NULL;
    int* core_st = NULL;
    const int* status = status_in;
    if (outliers) {
        oflag = malloc(sizeof(int) * (n > 0 ? n : 1));
        core_st = malloc(sizeof(int) * (n > 0 ? n : 1));
        orc_fmm_outliers(n, x, y, z, status_in, oflag);
        for (int i = 0; i < n; i++) core_st[i] = oflag[i] ? 0 : status_in[i];
        status = core_st;
    }
    Basis B, B1;
    basis_init(&B, p);
    basis_init(&B1, p + 1);
    const int nc = B.nc;
    double root[4];
    orc_fmm_root(n, x, y, z, status, root);
    Tree t;
    memset(&t, 0, sizeof(t));
    t.L = L; t.ring = ring; t.p = p;
    t.x0[0] = root[0]; t.x0[1] = root[1]; t.x0[2] = root[2]; t.side = root[3];
    t.nlev = L + 1;
    /* sort active sources by leaf key */
    KI* ki = malloc(sizeof(KI) * (n > 0 ? n : 1));
    int na = 0;
    for (int i = 0; i < n; i++)
        if (status[i] == 1 || status[i] == 2) {
            ki[na].k = orc_fmm_key(x[i], y[i], z[i], root, L);
            ki[na].i = i;
            na++;
        }
    qsort(ki, na, sizeof(KI), cmp_ki);
    t.nbox = calloc(t.nlev, sizeof(int));
    t.key = calloc(t.nlev, sizeof(uint64_t*));
    t.first = calloc(t.nlev, sizeof(int*));
    t.count = calloc(t.nlev, sizeof(int*));
    t.M = calloc(t.nlev, sizeof(double*));
    t.Lc = calloc(t.nlev, sizeof(double*));
    /* leaf level */
    {
        int nb = 0;
        for (int i = 0; i < na; i++) if (i == 0 || ki[i].k != ki[i - 1].k) nb++;
        t.nbox[L] = nb;
        t.key[L] = malloc(sizeof(uint64_t) * (nb + 1));
        t.first[L] = malloc(sizeof(int) * (nb + 1));
        t.count[L] = malloc(sizeof(int) * (nb + 1));
        int b = -1;
        for (int i = 0; i < na; i++) {
            if (i == 0 || ki[i].k != ki[i - 1].k) {
                b++;
                t.key[L][b] = ki[i].k;
                t.first[L][b] = i;
                t.count[L][b] = 0;
            }
            t.count[L][b]++;
        }
    }
    for (int l = L - 1; l >= 0; l--) {
        int nb = 0, nch = t.nbox[l + 1];
        for (int i = 0; i < nch; i++)
            if (i == 0 || (t.key[l + 1][i] >> 3) != (t.key[l + 1][i - 1] >> 3)) nb++;
        t.nbox[l] = nb;
        t.key[l] = malloc(sizeof(uint64_t) * (nb + 1));
        t.first[l] = malloc(sizeof(int) * (nb + 1));
        t.count[l] = malloc(sizeof(int) * (nb + 1));
        int b = -1;
        for (int i = 0; i < nch; i++) {
            uint64_t pk = t.key[l + 1][i] >> 3;
            if (i == 0 || pk != (t.key[l + 1][i - 1] >> 3)) {
                b++;
                t.key[l][b] = pk;
                t.first[l][b] = i;
                t.count[l][b] = 0;
            }
            t.count[l][b]++;
        }
    }
    for (int l = 0; l <= L; l++) {
        t.M[l] = calloc((size_t)t.nbox[l] * nc, sizeof(double));
        t.Lc[l] = calloc((size_t)t.nbox[l] * nc, sizeof(double));
    }
    const int nt = orc_get_threads();
    /* P2M */
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int b = 0; b < t.nbox[L]; b++) {
        double c[3], h;
        box_center(&t, L, t.key[L][b], c, &h);
        double* M = t.M[L] + (size_t)b * nc;
        for (int s = t.first[L][b]; s < t.first[L][b] + t.count[L][b]; s++) {
            int j = ki[s].i;
            double v[3] = {(x[j] - c[0]) / h, (y[j] - c[1]) / h, (z[j] - c[2]) / h};
            for (int k = 0; k < nc; k++)
                M[k] += q[j] * ipow(v[0], B.mi[k][0]) * ipow(v[1], B.mi[k][1]) * ipow(v[2], B.mi[k][2]);
        }
    }
    /* M2M */
    for (int l = L - 1; l >= 0; l--) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
        for (int b = 0; b < t.nbox[l]; b++) {
            double cp[3], hp;
            box_center(&t, l, t.key[l][b], cp, &hp);
            double* Mp = t.M[l] + (size_t)b * nc;
            for (int ch = t.first[l][b]; ch < t.first[l][b] + t.count[l][b]; ch++) {
                double cc[3], hc;
                box_center(&t, l + 1, t.key[l + 1][ch], cc, &hc);
                double d[3] = {(cc[0] - cp[0]) / hp, (cc[1] - cp[1]) / hp, (cc[2] - cp[2]) / hp};
                const double* Mc = t.M[l + 1] + (size_t)ch * nc;
                for (int a = 0; a < nc; a++) {
                    int* A = B.mi[a];
                    double acc = 0.0;
                    for (int k0 = 0; k0 <= A[0]; k0++)
                        for (int k1 = 0; k1 <= A[1]; k1++)
                            for (int k2 = 0; k2 <= A[2]; k2++) {
                                double w = binom(A[0], k0) * binom(A[1], k1) * binom(A[2], k2) *
                                           ipow(d[0], A[0] - k0) * ipow(d[1], A[1] - k1) *
                                           ipow(d[2], A[2] - k2) * ldexp(1.0, -(k0 + k1 + k2));
                                acc += w * Mc[B.idx[k0][k1][k2]];
                            }
                    Mp[a] += acc;
                }
            }
        }
    }
    /* M2L.  The weights C(a+b, a) (-1)^|a| and the index of T_{a+b} of every
     * (b, a) term are tabulated once; each is the same product as written in
     * the header, so the sums are unchanged. */
    double nm2l = 0, np2p = 0;
    const double eps2 = 6.4e-19;
    double* W = malloc(sizeof(double) * (size_t)nc * nc);
    int* TI = malloc(sizeof(int) * (size_t)nc * nc);
    for (int bb = 0; bb < nc; bb++) {
        int* Bm = B.mi[bb];
        for (int a = 0; a < nc; a++) {
            int* A = B.mi[a];
            int na_ = A[0] + A[1] + A[2], nb_ = Bm[0] + Bm[1] + Bm[2];
            double w = 0.0;
            int ti = 0;
            if (na_ + nb_ <= p) {
                w = binom(A[0] + Bm[0], A[0]) * binom(A[1] + Bm[1], A[1]) * binom(A[2] + Bm[2], A[2]);
                if (na_ & 1) w = -w;
                ti = B.idx[A[0] + Bm[0]][A[1] + Bm[1]][A[2] + Bm[2]];
            }
            W[(size_t)bb * nc + a] = w;
            TI[(size_t)bb * nc + a] = ti;
        }
    }
    for (int l = 1; l <= L && !treecode; l++) {
#pragma omp parallel num_threads(nt) reduction(+ : nm2l)
        {
        double* T = malloc(sizeof(double) * nc);
#pragma omp for schedule(dynamic, 8)
        for (int b = 0; b < t.nbox[l]; b++) {
            uint64_t key = t.key[l][b];
            int64_t ix[3];
            for (int d = 0; d < 3; d++) ix[d] = (int64_t)compact21(key >> d);
            double ca[3], h;
            box_center(&t, l, key, ca, &h);
            double* Lb = t.Lc[l] + (size_t)b * nc;
            int64_t px[3] = {ix[0] >> 1, ix[1] >> 1, ix[2] >> 1};
            int64_t nmax = ((int64_t)1 << (l - 1)) - 1;
            for (int64_t dz = -ring; dz <= ring; dz++)
                for (int64_t dy = -ring; dy <= ring; dy++)
                    for (int64_t dx = -ring; dx <= ring; dx++) {
                        int64_t qx = px[0] + dx, qy = px[1] + dy, qz = px[2] + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx > nmax || qy > nmax || qz > nmax) continue;
                        uint64_t pk = spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2;
                        int pb = find_box(&t, l - 1, pk);
                        if (pb < 0) continue;
                        for (int ch = t.first[l - 1][pb]; ch < t.first[l - 1][pb] + t.count[l - 1][pb]; ch++) {
                            uint64_t sk = t.key[l][ch];
                            int64_t sx = compact21(sk), sy = compact21(sk >> 1), sz = compact21(sk >> 2);
                            if (llabs(sx - ix[0]) <= ring && llabs(sy - ix[1]) <= ring &&
                                llabs(sz - ix[2]) <= ring)
                                continue;  /* near: handled below */
                            double cs[3], hs;
                            box_center(&t, l, sk, cs, &hs);
                            double R[3] = {(ca[0] - cs[0]) / h, (ca[1] - cs[1]) / h, (ca[2] - cs[2]) / h};
                            orc_fmm_taylor(p, R, eps2 / (h * h), T);
                            const double* Ms = t.M[l] + (size_t)ch * nc;
                            int64_t shell = llabs(sx - ix[0]);
                            if (llabs(sy - ix[1]) > shell) shell = llabs(sy - ix[1]);
                            if (llabs(sz - ix[2]) > shell) shell = llabs(sz - ix[2]);
                            const int pp = (shell < 8 && g_pshell[shell] > 0 && g_pshell[shell] < p)
                                               ? g_pshell[shell] : p;
                            for (int bb = 0; bb < nc; bb++) {
                                int* Bm = B.mi[bb];
                                int nb_ = Bm[0] + Bm[1] + Bm[2];
                                if (nb_ > pp) break;
                                const double* Wb = W + (size_t)bb * nc;
                                const int* Tb = TI + (size_t)bb * nc;
                                double acc = 0.0;
                                for (int a = 0; a < nc; a++) {
                                    int* A = B.mi[a];
                                    int na_ = A[0] + A[1] + A[2];
                                    if (na_ + nb_ > pp) break;
                                    acc += Wb[a] * T[Tb[a]] * Ms[a];
                                }
                                Lb[bb] += acc;
                            }
                            nm2l += 1;
                        }
                    }
        }
        free(T);
        }
    }
    free(W);
    free(TI);
    /* L2L */
    for (int l = 1; l < L && !treecode; l++) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
        for (int b = 0; b < t.nbox[l]; b++) {
            double cp[3], hp;
            box_center(&t, l, t.key[l][b], cp, &hp);
            const double* Lp = t.Lc[l] + (size_t)b * nc;
            for (int ch = t.first[l][b]; ch < t.first[l][b] + t.count[l][b]; ch++) {
                double cc[3], hc;
                box_center(&t, l + 1, t.key[l + 1][ch], cc, &hc);
                double e[3] = {(cc[0] - cp[0]) / hp, (cc[1] - cp[1]) / hp, (cc[2] - cp[2]) / hp};
                double* Lc = t.Lc[l + 1] + (size_t)ch * nc;
                for (int k = 0; k < nc; k++) {
                    int* K = B.mi[k];
                    double acc = 0.0;
                    for (int bb = 0; bb < nc; bb++) {
                        int* Bm = B.mi[bb];
                        if (Bm[0] < K[0] || Bm[1] < K[1] || Bm[2] < K[2]) continue;
                        acc += binom(Bm[0], K[0]) * binom(Bm[1], K[1]) * binom(Bm[2], K[2]) *
                               ipow(e[0], Bm[0] - K[0]) * ipow(e[1], Bm[1] - K[1]) *
                               ipow(e[2], Bm[2] - K[2]) * Lp[bb];
                    }
                    Lc[k] += 0.5 * ldexp(acc, -(K[0] + K[1] + K[2]));
                }
            }
        }
    }
    /* L2P + P2P */
    const double t_far = omp_get_wtime();
    double K = orc_k_const();
    for (int i = 0; i < 3 * n; i++) out3[i] = 0.0;
    const int stride = g_leaf_stride;
    const int nleaf_eval = (t.nbox[L] + stride - 1) / stride;
    double nm2p = 0;
#pragma omp parallel num_threads(nt) reduction(+ : np2p, nm2p)
    {
    double* T1 = malloc(sizeof(double) * B1.nc);
#pragma omp for schedule(dynamic, 8)
    for (int bs = 0; bs < nleaf_eval; bs++) {
        const int b = bs * stride;
        uint64_t key = t.key[L][b];
        double c[3], h;
        box_center(&t, L, key, c, &h);
        const double* Lb = t.Lc[L] + (size_t)b * nc;
        int64_t ix[3];
        for (int d = 0; d < 3; d++) ix[d] = (int64_t)compact21(key >> d);
        int64_t nmax = ((int64_t)1 << L) - 1;
        for (int s = t.first[L][b]; s < t.first[L][b] + t.count[L][b]; s++) {
            int i = ki[s].i;
            if (status[i] != 1) continue;
            double u[3] = {(x[i] - c[0]) / h, (y[i] - c[1]) / h, (z[i] - c[2]) / h};
            double g[3] = {0, 0, 0};
            for (int k = 1; k < nc; k++) {
                int* Km = B.mi[k];
                for (int d = 0; d < 3; d++) {
                    if (Km[d] == 0) continue;
                    double term = Km[d] * Lb[k];
                    for (int e = 0; e < 3; e++) term *= ipow(u[e], Km[e] - (e == d));
                    g[d] += term;
                }
            }
            /* phi = (1/h) sum L u^b -> grad_x = (1/h^2) grad_u;  P2P part as sum q (x-y)/r^3 */
            double ax = -g[0] / (h * h), ay = -g[1] / (h * h), az = -g[2] / (h * h);
            for (int64_t dz = -ring; dz <= ring; dz++)
                for (int64_t dy = -ring; dy <= ring; dy++)
                    for (int64_t dx = -ring; dx <= ring; dx++) {
                        int64_t qx = ix[0] + dx, qy = ix[1] + dy, qz = ix[2] + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx > nmax || qy > nmax || qz > nmax) continue;
                        uint64_t nk = spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2;
                        int nb = find_box(&t, L, nk);
                        if (nb < 0) continue;
                        for (int s2 = t.first[L][nb]; s2 < t.first[L][nb] + t.count[L][nb]; s2++) {
                            int j = ki[s2].i;
                            if (j == i) continue;
                            double ddx = x[i] - x[j], ddy = y[i] - y[j], ddz = z[i] - z[j];
                            double inv = 1.0 / sqrt(ddx * ddx + ddy * ddy + ddz * ddz + eps2);
                            double w = q[j] * inv * inv * inv;
                            ax += ddx * w; ay += ddy * w; az += ddz * w;
                            np2p += 1;
                        }
                    }
            if (treecode) {
                /* far field by M2P at every level (no local expansions) */
                double X[3] = {x[i], y[i], z[i]}, f[3] = {0, 0, 0};
                nm2p += m2p_point(&t, &B, &B1, ki, x, y, z, q, X, eps2, 0, T1, f);
                ax += f[0]; ay += f[1]; az += f[2];
            }
            double kc = (double)((long double)(K * q[i]) / orc_species_mass_ld(species[i]));
            out3[3 * i] = ax * kc;
            out3[3 * i + 1] = ay * kc;
            out3[3 * i + 2] = az * kc;
        }
    }
    free(T1);
    }
    if (outliers) {
        /* outlier targets: the tree's field by M2P and their near leaf cells
         * (m2outlier_p, p2outlier_p); every target: the outliers' field directly
         * (outlier_p2p, fmm_cpukernel.cpp:2230-2280) */
        int nout = 0;
        for (int i = 0; i < n; i++) nout += oflag[i];
        int* ol = malloc(sizeof(int) * (nout > 0 ? nout : 1));
        for (int i = 0, k = 0; i < n; i++)
            if (oflag[i]) ol[k++] = i;
#pragma omp parallel num_threads(nt) reduction(+ : nm2p)
        {
        double* T1 = malloc(sizeof(double) * B1.nc);
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < n; i++) {
            if (status_in[i] != 1) continue;
            double f[3] = {0, 0, 0};
            if (oflag[i]) {
                double X[3] = {x[i], y[i], z[i]};
                nm2p += m2p_point(&t, &B, &B1, ki, x, y, z, q, X, eps2, 1, T1, f);
            }
            for (int k = 0; k < nout; k++) {
                int j = ol[k];
                if (j == i) continue;
                double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
                double inv = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + eps2);
                double wq = q[j] * inv * inv * inv;
                f[0] += dx * wq; f[1] += dy * wq; f[2] += dz * wq;
            }
            double kc = (double)((long double)(K * q[i]) / orc_species_mass_ld(species[i]));
            out3[3 * i] += f[0] * kc;
            out3[3 * i + 1] += f[1] * kc;
            out3[3 * i + 2] += f[2] * kc;
        }
        free(T1);
        }
        free(ol);
    }
    const double t_end = omp_get_wtime();
    if (stats) {
        stats[0] = t.nbox[L];
        stats[1] = np2p;
        stats[2] = nm2l;
        stats[3] = t_far - t_start;
        stats[4] = t_end - t_far;
        stats[5] = nleaf_eval;
        stats[6] = nm2p;
    }
    for (int l = 0; l <= L; l++) {
        free(t.key[l]); free(t.first[l]); free(t.count[l]); free(t.M[l]); free(t.Lc[l]);
    }
    free(t.nbox); free(t.key); free(t.first); free(t.count); free(t.M); free(t.Lc);
    free(ki);
    free(B.mi);
    free(B1.mi);
    free(oflag);
    free(core_st);
}

/* Interaction counts of the FMM tree definition above, without expansions:
 * out[0] = M2L box pairs over all levels, out[1] = P2P particle pairs
 * (self pairs included), out[2] = leaf boxes. */
static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}
static int find_u64(const uint64_t* k, int nb, uint64_t key) {
    int lo = 0, hi = nb - 1;
    while (lo <= hi) {
        int mid = (lo + hi) >> 1;
        if (k[mid] == key) return mid;
        if (k[mid] < key) lo = mid + 1; else hi = mid - 1;
    }
    return -1;
}
void orc_fmm_counts(int n, const double* x, const double* y, const double* z, const int* status,
                    int L, int ring, double* out) {
    double root[4];
    orc_fmm_root(n, x, y, z, status, root);
    uint64_t* k = malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    int na = 0;
    for (int i = 0; i < n; i++)
        if (status[i] == 1 || status[i] == 2) k[na++] = orc_fmm_key(x[i], y[i], z[i], root, L);
    qsort(k, na, sizeof(uint64_t), cmp_u64);
    uint64_t** bk = calloc(L + 1, sizeof(uint64_t*));
    int** bc = calloc(L + 1, sizeof(int*));
    int* nb = calloc(L + 1, sizeof(int));
    for (int l = L; l >= 0; l--) {
        int sh = 3 * (L - l);
        bk[l] = malloc(sizeof(uint64_t) * (na + 1));
        bc[l] = malloc(sizeof(int) * (na + 1));
        int m = 0;
        for (int i = 0; i < na; i++) {
            uint64_t v = k[i] >> sh;
            if (m == 0 || bk[l][m - 1] != v) { bk[l][m] = v; bc[l][m] = 0; m++; }
            bc[l][m - 1]++;
        }
        nb[l] = m;
    }
    double m2l = 0, p2p = 0;
    for (int l = 1; l <= L; l++) {
        int64_t pmax = ((int64_t)1 << (l - 1)) - 1;
        for (int b = 0; b < nb[l]; b++) {
            int64_t ix = compact21(bk[l][b]), iy = compact21(bk[l][b] >> 1), iz = compact21(bk[l][b] >> 2);
            for (int dz = -ring; dz <= ring; dz++)
                for (int dy = -ring; dy <= ring; dy++)
                    for (int dx = -ring; dx <= ring; dx++) {
                        int64_t qx = (ix >> 1) + dx, qy = (iy >> 1) + dy, qz = (iz >> 1) + dz;
                        if (qx < 0 || qy < 0 || qz < 0 || qx > pmax || qy > pmax || qz > pmax) continue;
                        if (find_u64(bk[l - 1], nb[l - 1], spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2) < 0)
                            continue;
                        for (int o = 0; o < 8; o++) {
                            int64_t cx = 2 * qx + (o & 1), cy = 2 * qy + ((o >> 1) & 1), cz = 2 * qz + ((o >> 2) & 1);
                            if (llabs(cx - ix) <= ring && llabs(cy - iy) <= ring && llabs(cz - iz) <= ring) continue;
                            if (find_u64(bk[l], nb[l], spread21(cx) | spread21(cy) << 1 | spread21(cz) << 2) >= 0)
                                m2l += 1;
                        }
                    }
        }
    }
    int64_t cmax = ((int64_t)1 << L) - 1;
    for (int b = 0; b < nb[L]; b++) {
        int64_t ix = compact21(bk[L][b]), iy = compact21(bk[L][b] >> 1), iz = compact21(bk[L][b] >> 2);
        double s = 0;
        for (int dz = -ring; dz <= ring; dz++)
            for (int dy = -ring; dy <= ring; dy++)
                for (int dx = -ring; dx <= ring; dx++) {
                    int64_t qx = ix + dx, qy = iy + dy, qz = iz + dz;
                    if (qx < 0 || qy < 0 || qz < 0 || qx > cmax || qy > cmax || qz > cmax) continue;
                    int j = find_u64(bk[L], nb[L], spread21(qx) | spread21(qy) << 1 | spread21(qz) << 2);
                    if (j >= 0) s += bc[L][j];
                }
        p2p += s * bc[L][b];
    }
    out[0] = m2l; out[1] = p2p; out[2] = nb[L];
    for (int l = 0; l <= L; l++) { free(bk[l]); free(bc[l]); }
    free(bk); free(bc); free(nb); free(k);
}
