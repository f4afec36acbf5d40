/*
 * Oracle (test infrastructure only): CPU restatement of the reference's
 * Barnes-Hut force calculation, used to check the GPU engine.  Nothing in the
 * product path links or calls it.
 *
 * Root layout (the grid of root cells):
 *   - DomainBoundsFunction::computeDomainBounds, domain_bounds_functions.cpp:121-316
 *     (bounds seeded with the first ion of the range whatever its status;
 *     cubic cells: the smallest non-zero extent is the first cell size, the
 *     other axes are cut into floor/ceil(extent / size) cells, the size is
 *     doubled until the average cell holds >= FMMConstant::minAvgPPB = 100
 *     ions (constants.h:80), then scaled by 1 + domainEps / n with
 *     domainEps = 0.01f (domain_bounds_functions.h:39) and centred;
 *     rectangular-prism cells: one cell of the bounding box; forced cell
 *     counts from setBoxDimVals)
 *   - BarnesHutForceCalculation::computeWorldLimits,
 *     barnes_hut_force_calculation.cpp:456-522 (cells added until the grid
 *     covers every source; cubic grids merged 2x2x2 while the smallest count
 *     is >= 2)
 *   - the chain of root cells, :604-632 (x outermost, z innermost, linked as
 *     siblings) and their locations roots + i * dims, :263-277
 * The outlier toggle (setOutlierToggle) is off, as in the reference's default
 * (force_calculation_methods.h:316); with it off every outlier flag is 0.
 *
 * Tree and walk, barnes_hut_force_calculation.cpp:
 *   - generateTree :258-331 (ions of status 1/2 inserted in index order; the
 *     root cell from (pos - roots) / dims truncated to int, the last cell of
 *     the chain when no cell matches)
 *   - locateSubcell :149-216 (strict float compares against location +
 *     dimensions / 2), addToCell :224-253 (split until one ion per cell; all
 *     8 children exist; the split cell keeps its stale index)
 *   - computeCellProperties :337-390 (net charge, sum |q|, centre of charge
 *     weighted by |net charge of child| over sum |q|, geometric centre when
 *     that sum is 0, delta = |centre of charge - geometric centre|)
 *   - computeEffectiveDistance :135-141, computeForceFromCell :399-409,
 *     computeForceFromTree :417-450 (stackless threaded walk over the whole
 *     chain; MAC theta (d - delta) > repLength)
 *   - barnesHutKernelCalculation :539-569 (accel *= K q / m in double,
 *     rounded to float per component)
 * and cell.cpp:7-131 (children at parent + bit * dims / 2, repLength = the
 * largest dimension), the rotational variant
 * barnes_hut_force_calculation_rotational.cpp:28-135 (four quadrants, target
 * rotated forward, contributions rotated back, self excluded in quadrant 0
 * only, each quadrant's float result added to the float total) and the
 * translational variant barnes_hut_force_calculation_translational.cpp:190-280
 * (target shifted by each lattice image, self excluded at the central image
 * only).
 *
 * One case differs, shared with the GPU engine: ions that agree down to 42
 * levels below their root cell (closer than 2^-42 of the root cell, i.e.
 * coincident ions) stay together in one bucket leaf.  The reference's
 * addToCell recursion does not terminate for coincident ions.
 */
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BH_MAX_DEPTH 42
#define BH_MIN_AVG_PPB 100          /* FMMConstant::minAvgPPB, constants.h:80 */
#define BH_DOMAIN_EPS 0.01f          /* domain_bounds_functions.h:39 */
#define BH_SOFTENING 6.4e-19         /* PhysicalConstant::SOFTENING, constants.h:31 */

/* float -> int as x86-64's cvttss2si does it (the reference's platform):
 * truncation, INT_MIN for NaN and out-of-range values. */
static int bh_f2i(float v) {
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int)v : INT_MIN;
}

/* ---------------------------------------------------------- root layout -- */

typedef struct {
    float roots[3];  /* location of the first root cell */
    float dims[3];   /* dimensions of every root cell */
    int nd[3];       /* domainDimVals: cells per axis */
    int nsrc;        /* ions of status 1/2 in the range */
} bh_layout;

/* DomainBoundsFunction::computeDomainBounds (no outliers): the corner of the
 * first root cell (boxMin), the cell dimensions and the cells per axis.
 * toggle: 0 cubic cells, 1 rectangular-prism cells; forced: NULL or the
 * cell counts of setBoxDimVals.  minO / maxO: the bounds of the sources
 * seeded with the first ion of the range.  Returns the number of sources. */
static int domain_bounds(const float* pos4, const float* acc4, int i0, int i1, int toggle, int rot,
                         const int* forced, float boxMin[3], float dims[3], int bdv[3],
                         float minO[3], float maxO[3]) {
    const float* seed = pos4 + 4 * (size_t)i0;
    float rootmax[3];
    for (int c = 0; c < 3; c++) boxMin[c] = rootmax[c] = minO[c] = maxO[c] = seed[c];
    if (rot) boxMin[0] = boxMin[1] = 0.0f; /* computeDomainBoundsRotations: x, y minima not taken */
    int npart = 0;
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        const float* p = pos4 + 4 * (size_t)i;
        for (int c = 0; c < 3; c++) {
            if (p[c] < boxMin[c] && (!rot || c == 2)) boxMin[c] = p[c];
            if (p[c] > rootmax[c]) rootmax[c] = p[c];
            if (p[c] < minO[c]) minO[c] = p[c];
            if (p[c] > maxO[c]) maxO[c] = p[c];
        }
        npart++;
    }
    const float subtr = 0.0f;
    if (rot) {
        /* a square about the z axis: x, y in [-R, R], R the larger of the x and y maxima
         * (domain_bounds_functions.cpp:393-403) */
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
        float dimDiff[3] = {dims[0], dims[1], dims[2]};
        int dimInd[3] = {0, 1, 2};
        /* insertion sort, smallest first (ties keep axis order) */
        for (int i = 0; i < 3; i++) {
            int j = i;
            while (j > 0 && dimDiff[j - 1] > dimDiff[j]) {
                const float td = dimDiff[j - 1];
                const int ti = dimInd[j - 1];
                dimDiff[j - 1] = dimDiff[j];
                dimInd[j - 1] = dimInd[j];
                dimDiff[j] = td;
                dimInd[j] = ti;
                j--;
            }
        }
        float rbs = 0.0f;
        int i = -1;
        do {
            i++;
            rbs = dimDiff[i];
            bdv[dimInd[i]] = 1;
        } while (rbs == 0.0f && i < 2);
        const float soft = (float)BH_SOFTENING;
        rbs = (rbs < soft) ? soft : rbs; /* std::max(rootBoxSize, (float)SOFTENING) */
        int found = 0;
        do {
            for (int j = i + 1; j < 3; j++) {
                const float ratio = dimDiff[j] / rbs;
                if ((double)ratio - floor((double)ratio) >= 0.5) {
                    bdv[dimInd[j]] = (int)ceil((double)ratio);
                } else {
                    if (ratio < 2.0f) {
                        bdv[dimInd[j]] = (j == 0) ? 1 : bdv[dimInd[j - 1]];
                    } else {
                        bdv[dimInd[j]] = (int)floor((double)ratio);
                    }
                    if ((double)ratio > 0.5 || j == 2) rbs = dimDiff[j] / (float)bdv[dimInd[j]];
                }
            }
            const int ndom = bdv[0] * bdv[1] * bdv[2];
            const int avg = (npart - 0 + ndom - 1) / ndom;
            if (rbs == dimDiff[2] || avg >= BH_MIN_AVG_PPB || rbs == soft) found = 1;
            else rbs *= 2.0f;
        } while (!found);
        if (forced) {
            for (int c = 0; c < 3; c++) bdv[c] = forced[c];
            rbs = dims[0] / (float)bdv[0];
            const float r1 = dims[1] / (float)bdv[1];
            rbs = (rbs < r1) ? r1 : rbs;
            const float r2 = dims[2] / (float)bdv[2];
            rbs = (rbs < r2) ? r2 : rbs;
            rbs = (rbs < soft) ? soft : rbs;
            if (bdv[0] >= bdv[1] && bdv[0] >= bdv[2]) dimInd[2] = 0;
            else if (bdv[1] >= bdv[0] && bdv[1] >= bdv[2]) dimInd[2] = 1;
            else if (bdv[2] >= bdv[0] && bdv[2] >= bdv[1]) dimInd[2] = 2;
        }
        rbs *= (1.0f + BH_DOMAIN_EPS / (float)bdv[dimInd[2]]);
        for (int c = 0; c < 3; c++) {
            boxMin[c] = ((rootmax[c] + subtr) + boxMin[c] - rbs * (float)bdv[c]) / 2.0f;
            dims[c] = rbs;
        }
    }
    return npart;
}

/* computeDomainBounds followed by computeWorldLimits. */
static void bh_make_layout(const float* pos4, const float* acc4, int i0, int i1, int toggle,
                           const int* forced, bh_layout* L) {
    float boxMin[3], dims[3], minO[3], maxO[3];
    int bdv[3];
    const int npart = domain_bounds(pos4, acc4, i0, i1, toggle, 0, forced, boxMin, dims, bdv, minO, maxO);
    /* computeWorldLimits */
    float roots[3] = {boxMin[0], boxMin[1], boxMin[2]};
    for (int c = 0; c < 3; c++) {
        while (roots[c] > minO[c]) {
            roots[c] -= dims[c];
            bdv[c]++;
        }
    }
    for (int c = 0; c < 3; c++)
        while (roots[c] + (float)bdv[c] * dims[c] < maxO[c]) bdv[c]++;
    if (toggle == 0) {
        float rbs = dims[0];
        int dc = bdv[0] < bdv[1] ? bdv[0] : bdv[1];
        dc = dc < bdv[2] ? dc : bdv[2];
        while (dc / 2 > 0) {
            dc = (dc + 1) / 2;
            rbs *= 2.0f;
            for (int c = 0; c < 3; c++) bdv[c] = (bdv[c] + 1) / 2;
        }
        for (int c = 0; c < 3; c++) dims[c] = rbs;
    }
    for (int c = 0; c < 3; c++) {
        L->roots[c] = roots[c];
        L->dims[c] = dims[c];
        L->nd[c] = bdv[c];
    }
    L->nsrc = npart;
}

/* ------------------------------------------------------------------ tree -- */

typedef struct {
    int first_child;  /* -1: leaf */
    int next;         /* threaded successor (-1 ends the walk) */
    int index;        /* ion of a leaf (-1: empty) */
    int depth;        /* levels below the root cell */
    float lx, ly, lz;     /* location (corner) */
    float dx, dy, dz;     /* dimensions */
    float rep;            /* repLength: largest dimension */
    float charge, uq;
    float cx, cy, cz;
    float delta;
} bh_cell;

typedef struct {
    bh_cell* c;
    int n, cap;
    int* ion_next;  /* bucket chains: next ion in the same depth-42 leaf */
} bh_tree;

static float bh_max3(float a, float b, float c) {
    /* Cell::Cell sorts the three dimensions ascending and keeps the last */
    float m = a;
    if (b > m) m = b;
    if (c > m) m = c;
    return m;
}

static int bh_new_cell(bh_tree* t, float lx, float ly, float lz, float dx, float dy, float dz,
                       int depth) {
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 1024;
        t->c = (bh_cell*)realloc(t->c, sizeof(bh_cell) * (size_t)t->cap);
    }
    bh_cell* c = &t->c[t->n];
    memset(c, 0, sizeof(*c));
    c->first_child = -1;
    c->next = -1;
    c->index = -1;
    c->depth = depth;
    c->lx = lx;
    c->ly = ly;
    c->lz = lz;
    c->dx = dx;
    c->dy = dy;
    c->dz = dz;
    c->rep = bh_max3(dx, dy, dz);
    return t->n++;
}

/* locateSubcell: strict compares against location + dimensions / 2 */
static int bh_subcell(const bh_cell* c, const float* p) {
    const float cx = c->lx + c->dx / 2.0f, cy = c->ly + c->dy / 2.0f, cz = c->lz + c->dz / 2.0f;
    return (p[0] > cx) + ((p[1] > cy) << 1) + ((p[2] > cz) << 2);
}

/* generateSubcells + setLocationOfSubcells: 8 children, threaded, the last
 * one continuing with the split cell's successor */
static void bh_split(bh_tree* t, int ci) {
    const float hx = t->c[ci].dx / 2.0f, hy = t->c[ci].dy / 2.0f, hz = t->c[ci].dz / 2.0f;
    int first = -1;
    for (int i = 0; i < 8; i++) {
        const bh_cell* p = &t->c[ci];
        float lx = p->lx, ly = p->ly, lz = p->lz;
        lx += (float)(i & 1) * hx;
        ly += (float)((i & 2) >> 1) * hy;
        lz += (float)((i & 4) >> 2) * hz;
        const int k = bh_new_cell(t, lx, ly, lz, hx, hy, hz, t->c[ci].depth + 1);
        if (i == 0) first = k;
    }
    for (int i = 0; i < 7; i++) t->c[first + i].next = first + i + 1;
    t->c[first + 7].next = t->c[ci].next;
    t->c[ci].first_child = first;
}

/* addToCell, with the recursion unrolled */
static void bh_add(bh_tree* t, const float* pos4, int ci, int ion) {
    for (;;) {
        if (t->c[ci].index == -1) {
            t->c[ci].index = ion;
            return;
        }
        if (t->c[ci].depth >= BH_MAX_DEPTH) { /* bucket: append to the chain */
            int j = t->c[ci].index;
            while (t->ion_next[j] >= 0) j = t->ion_next[j];
            t->ion_next[j] = ion;
            return;
        }
        const int old = t->c[ci].index;
        bh_split(t, ci);
        const int s1 = bh_subcell(&t->c[ci], pos4 + 4 * (size_t)old);
        t->c[t->c[ci].first_child + s1].index = old;
        const int s2 = bh_subcell(&t->c[ci], pos4 + 4 * (size_t)ion);
        if (s1 != s2) {
            t->c[t->c[ci].first_child + s2].index = ion;
            return;
        }
        ci = t->c[ci].first_child + s1;
    }
}

static void bh_geo_centre(const bh_cell* c, float* g) {
    g[0] = c->lx + c->dx / 2.0f;
    g[1] = c->ly + c->dy / 2.0f;
    g[2] = c->lz + c->dz / 2.0f;
}

static void bh_props(bh_tree* t, const float* pos4, int ci) {
    bh_cell* c = &t->c[ci];
    if (c->first_child < 0) {
        if (c->index >= 0) {
            if (t->ion_next[c->index] < 0) {
                const float* p = pos4 + 4 * (size_t)c->index;
                c->charge = p[3];
                c->uq = fabsf(p[3]);
                c->cx = p[0];
                c->cy = p[1];
                c->cz = p[2];
            } else { /* bucket: the sums of a cell with its ions as children */
                float tx = 0.0f, ty = 0.0f, tz = 0.0f;
                for (int j = c->index; j >= 0; j = t->ion_next[j]) {
                    const float* p = pos4 + 4 * (size_t)j;
                    c->charge += p[3];
                    c->uq += fabsf(p[3]);
                    tx += p[0] * fabsf(p[3]);
                    ty += p[1] * fabsf(p[3]);
                    tz += p[2] * fabsf(p[3]);
                }
                if (c->uq != 0.0f) {
                    c->cx = tx / c->uq;
                    c->cy = ty / c->uq;
                    c->cz = tz / c->uq;
                } else {
                    float g[3];
                    bh_geo_centre(c, g);
                    c->cx = g[0];
                    c->cy = g[1];
                    c->cz = g[2];
                }
            }
        }
        /* an empty leaf keeps charge 0 and centre of charge 0 (Cell::Cell) */
    } else {
        float tx = 0.0f, ty = 0.0f, tz = 0.0f;
        for (int i = 0; i < 8; i++) {
            bh_props(t, pos4, t->c[ci].first_child + i);
            c = &t->c[ci];
            const bh_cell* s = &t->c[c->first_child + i];
            c->charge += s->charge;
            c->uq += s->uq;
            tx += s->cx * fabsf(s->charge);
            ty += s->cy * fabsf(s->charge);
            tz += s->cz * fabsf(s->charge);
        }
        if (c->uq != 0.0f) {
            c->cx = tx / c->uq;
            c->cy = ty / c->uq;
            c->cz = tz / c->uq;
        } else {
            float g[3];
            bh_geo_centre(c, g);
            c->cx = g[0];
            c->cy = g[1];
            c->cz = g[2];
        }
    }
    float g[3];
    bh_geo_centre(c, g);
    const float dx = c->cx - g[0], dy = c->cy - g[1], dz = c->cz - g[2];
    c->delta = sqrtf(dx * dx + dy * dy + dz * dz);
}

/* computeEffectiveDistance: float squares, double softening and sqrt */
static float bh_dist(const float* p, float cx, float cy, float cz, float* rx, float* ry, float* rz) {
    *rx = p[0] - cx;
    *ry = p[1] - cy;
    *rz = p[2] - cz;
    const float s = *rx * *rx + *ry * *ry + *rz * *rz;
    return (float)sqrt((double)s + BH_SOFTENING);
}

static void bh_accum(float f, float rx, float ry, float rz, int cb, int sb, float* acc) {
    acc[0] += f * ((float)cb * rx - (float)sb * ry);
    acc[1] += f * ((float)sb * rx + (float)cb * ry);
    acc[2] += f * rz;
}

/* One walk for target position p (already transformed) starting at the first
 * root cell; (cb, sb) rotate the contributions back (1, 0 for none); the ion
 * `self` is skipped; counts[0] += accepted cells, counts[1] += ion terms,
 * counts[2] += opened cells. */
static void bh_walk(const bh_tree* t, const float* pos4, const float* p, int self, float theta,
                    int cb, int sb, float* acc, long long* counts) {
    int ci = 0;
    while (ci >= 0) {
        const bh_cell* c = &t->c[ci];
        if (c->first_child < 0) {
            if (c->index >= 0) {
                if (t->ion_next[c->index] < 0) {
                    if (c->index != self) {
                        float rx, ry, rz;
                        const float d = bh_dist(p, c->cx, c->cy, c->cz, &rx, &ry, &rz);
                        bh_accum(c->charge / (d * d * d), rx, ry, rz, cb, sb, acc);
                        counts[1]++;
                    }
                } else {
                    for (int j = c->index; j >= 0; j = t->ion_next[j]) {
                        if (j == self) continue;
                        const float* q = pos4 + 4 * (size_t)j;
                        float rx, ry, rz;
                        const float d = bh_dist(p, q[0], q[1], q[2], &rx, &ry, &rz);
                        bh_accum(q[3] / (d * d * d), rx, ry, rz, cb, sb, acc);
                        counts[1]++;
                    }
                }
            }
            ci = c->next;
        } else {
            float rx, ry, rz;
            const float d = bh_dist(p, c->cx, c->cy, c->cz, &rx, &ry, &rz);
            if (theta * (d - c->delta) > c->rep) {
                bh_accum(c->charge / (d * d * d), rx, ry, rz, cb, sb, acc);
                counts[0]++;
                ci = c->next;
            } else {
                counts[2]++;
                ci = c->first_child;
            }
        }
    }
}

/* quadrantRotationFunction, auxiliary_functions.h:55-57 */
static int bh_qrf(int n) {
    const int oe = (n & 1) ? -1 : 1;
    const int a = 1 - oe;
    const int sg = (a > 0) - (a < 0);
    const int oe2 = (((n - 1) >> 1) & 1) ? -1 : 1;
    return sg * oe2;
}

static double bh_accel_const(float q, int species) {
    long double m;
    switch (species) {
    case 0: m = 111.0L; break;
    case 1: m = 309.0L; break;
    case 2: m = 507.0L; break;
    case 4: m = 198.0L; break;
    case 5: m = 87.0L; break;
    case 6: m = 285.0L; break;
    case 7: m = 483.0L; break;
    default: m = 0.0L; break;
    }
    extern double orc_k_const(void);
    const double kq = orc_k_const() * (double)q;
    return (double)((long double)kq / m);
}

/* Root layout of the sources in [i0, i1): roots3, dims3 (float), nd3 (cells
 * per axis).  toggle 0 cubic, 1 rectangular prism; forced3 NULL or the cell
 * counts.  Returns the number of sources. */
int orc_bh_layout(const float* pos4, const float* acc4, int i0, int i1, int toggle,
                  const int* forced3, float* roots3, float* dims3, int* nd3) {
    bh_layout L;
    if (i1 <= i0) return 0;
    bh_make_layout(pos4, acc4, i0, i1, toggle, forced3, &L);
    for (int c = 0; c < 3; c++) {
        roots3[c] = L.roots[c];
        dims3[c] = L.dims[c];
        nd3[c] = L.nd[c];
    }
    return L.nsrc;
}

/* computeDomainBounds alone, as MultipoleMethodForceCalculation::setDomainSize
 * calls it (multipole_method_force_calculation.cpp:648-668): cubic cells, no
 * outliers.  rot != 0: computeDomainBoundsRotations (domain_bounds_functions.cpp:
 * 317-500, called by setDomainSizeRotations, multipole_method_force_calculation_
 * rotational.cpp:333-350), whose x, y extent is the square [-R, R]^2 about the
 * z axis.  Returns the number of sources; box_min3 / dims3 / nd3 as the
 * reference's boxMin, dims and boxDimVals. */
int orc_domain_bounds(const float* pos4, const float* acc4, int i0, int i1, const int* forced3, int rot,
                      float* box_min3, float* dims3, int* nd3) {
    float minO[3], maxO[3];
    if (i1 <= i0) return 0;
    return domain_bounds(pos4, acc4, i0, i1, 0, rot, forced3, box_min3, dims3, nd3, minO, maxO);
}

/* The root cell an ion is placed in (generateTree :279-310), as a chain index;
 * -1 when no cell matches (the reference then continues in the last one). */
static int bh_root_of(const bh_layout* L, const float* p) {
    const int di = bh_f2i((p[0] - L->roots[0]) / L->dims[0]);
    const int dj = bh_f2i((p[1] - L->roots[1]) / L->dims[1]);
    const int dk = bh_f2i((p[2] - L->roots[2]) / L->dims[2]);
    if (di < 0 || di >= L->nd[0] || dj < 0 || dj >= L->nd[1] || dk < 0 || dk >= L->nd[2]) return -1;
    return (di * L->nd[1] + dj) * L->nd[2] + dk;
}

/*
 * Barnes-Hut accelerations of targets (status 1) in [i0, i1) from sources
 * (status 1 or 2) in the same range.  sym: 0 none, 1 rotational (four
 * quadrants), 2 translational (nimg lattice points times domain, in the
 * given order).  toggle / forced3: root layout options as for orc_bh_layout.
 * out3[3 (i - i0) + c] = the float acceleration the reference accumulates
 * (zero start), as double; zero for non-targets.  info (may be NULL):
 * [0] cells of the tree including empty ones, [1] root cells, [2] deepest
 * level, [3] bucket leaves, [4] accepted cells, [5] ion terms, [6] ions
 * placed in the last root cell because no cell matched, [7] cells that hold
 * at least one ion, [8] opened cells (the walk descended into them).
 */
int orc_bh_ref2(const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1,
                float theta, int sym, int nimg, const int* lattice, const double* domain, int toggle,
                const int* forced3, double* out3, long long* info) {
    (void)n;
    long long inf[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = i0; i < i1; i++) out3[3 * (size_t)(i - i0)] = out3[3 * (size_t)(i - i0) + 1] =
                                        out3[3 * (size_t)(i - i0) + 2] = 0.0;
    if (i1 <= i0) {
        if (info) memcpy(info, inf, sizeof(inf));
        return 0;
    }
    bh_layout L;
    bh_make_layout(pos4, acc4, i0, i1, toggle, forced3, &L);
    bh_tree t = {0};
    t.ion_next = (int*)malloc(sizeof(int) * (size_t)i1);
    for (int i = 0; i < i1; i++) t.ion_next[i] = -1;
    const int nroot = L.nd[0] * L.nd[1] * L.nd[2];
    for (int i = 0; i < L.nd[0]; i++)
        for (int j = 0; j < L.nd[1]; j++)
            for (int k = 0; k < L.nd[2]; k++) {
                const int r = bh_new_cell(&t, L.roots[0] + (float)i * L.dims[0],
                                          L.roots[1] + (float)j * L.dims[1],
                                          L.roots[2] + (float)k * L.dims[2], L.dims[0], L.dims[1],
                                          L.dims[2], 0);
                if (r > 0) t.c[r - 1].next = r;
            }
    if (nroot > 0) {
        t.c[0].lx = L.roots[0];
        t.c[0].ly = L.roots[1];
        t.c[0].lz = L.roots[2];
    }
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        const float* p = pos4 + 4 * (size_t)i;
        int ci = bh_root_of(&L, p);
        if (ci < 0) {
            ci = nroot - 1;
            inf[6]++;
        }
        while (t.c[ci].first_child >= 0) ci = t.c[ci].first_child + bh_subcell(&t.c[ci], p);
        bh_add(&t, pos4, ci, i);
    }
    for (int r = 0; r < nroot; r++) bh_props(&t, pos4, r);
    for (int c = 0; c < t.n; c++) {
        if (t.c[c].depth > inf[2]) inf[2] = t.c[c].depth;
        if (t.c[c].first_child < 0 && t.c[c].index >= 0 && t.ion_next[t.c[c].index] >= 0) inf[3]++;
        if (t.c[c].first_child >= 0 || t.c[c].index >= 0) inf[7]++;
    }
    for (int i = i0; i < i1; i++) {
        double* o = out3 + 3 * (size_t)(i - i0);
        if (acc4[4 * (size_t)i + 3] != 1.0f) continue;
        const float* p = pos4 + 4 * (size_t)i;
        const double kc = bh_accel_const(p[3], (int)vel4[4 * (size_t)i + 3]);
        float tot[3] = {0.0f, 0.0f, 0.0f};
        const int nrep = sym == 1 ? 4 : (sym == 2 ? nimg : 1);
        for (int k = 0; k < nrep; k++) {
            float tp[3] = {p[0], p[1], p[2]};
            int cb = 1, sb = 0, self = i;
            if (sym == 1) {
                const int cf = bh_qrf(5 - k), sf = bh_qrf(4 - k);
                cb = bh_qrf(k + 1);
                sb = bh_qrf(k);
                tp[0] = (float)cf * p[0] - (float)sf * p[1];
                tp[1] = (float)sf * p[0] + (float)cf * p[1];
                if (k != 0) self = -1;
            } else if (sym == 2) {
                const int* l = lattice + 3 * (size_t)k;
                for (int c = 0; c < 3; c++) tp[c] = p[c] - (float)l[c] * (float)domain[c];
                if (l[0] != 0 || l[1] != 0 || l[2] != 0) self = -1;
            }
            float acc[3] = {0.0f, 0.0f, 0.0f};
            long long wc[3] = {0, 0, 0};
            bh_walk(&t, pos4, tp, self, theta, cb, sb, acc, wc);
            inf[4] += wc[0];
            inf[5] += wc[1];
            inf[8] += wc[2];
            for (int c = 0; c < 3; c++) tot[c] += (float)((double)acc[c] * kc);
        }
        for (int c = 0; c < 3; c++) o[c] = (double)tot[c];
    }
    inf[0] = t.n;
    inf[1] = nroot;
    if (info) memcpy(info, inf, sizeof(inf));
    free(t.c);
    free(t.ion_next);
    return (int)t.n;
}

/* Default layout (cubic cells, no forced counts): the reference's defaults. */
int orc_bh_ref(const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1,
               float theta, int sym, int nimg, const int* lattice, const double* domain,
               double* out3) {
    return orc_bh_ref2(pos4, vel4, acc4, n, i0, i1, theta, sym, nimg, lattice, domain, 0, NULL, out3,
                       NULL);
}
