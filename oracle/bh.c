/*
 * Oracle (test infrastructure only): CPU restatement of the reference's
 * Barnes-Hut force calculation, used to check the GPU engine.
 *
 * Follows barnes_hut_force_calculation.cpp:
 *   - computeEffectiveDistance :135-141 (float distance, double softening, sqrt)
 *   - locateSubcell :149-216 (strict float compares against the cell centre)
 *   - addToCell :224-253 (split until one ion per cell; all 8 children exist)
 *   - generateTree :258-331 (ions of status 1/2 inserted in index order)
 *   - computeCellProperties :337-390 (net charge, sum |q|, centre of charge
 *     weighted by |net charge of child| over sum |q|, delta)
 *   - computeForceFromCell :399-409, computeForceFromTree :417-450 (stackless
 *     threaded walk; MAC theta (d - delta) > l)
 *   - barnesHutKernelCalculation :539-569 (accel *= K q / m in double,
 *     rounded to float per component)
 * and cell.cpp:7-131 (Cell geometry: children located at parent + bit *
 * dims / 2, repLength = largest dimension), the rotational variant
 * barnes_hut_force_calculation_rotational.cpp:28-135 (four quadrants, target
 * rotated forward, contributions rotated back, self excluded in quadrant 0
 * only) and the translational variant
 * barnes_hut_force_calculation_translational.cpp:190-280 (target shifted by
 * each lattice image, self excluded at the central image only).
 *
 * One difference, shared with the GPU engine and documented there: the root
 * is a single cube around the bounding box of the sources (side = largest
 * extent x (1 + 2^-20), centred on the box), where the reference lays a grid
 * of root cells from DomainBoundsFunction::computeDomainBounds.  A cell at
 * depth 21 is not split further (its ions form one bucket leaf); the
 * reference would keep splitting.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define BH_MAX_DEPTH 21

typedef struct {
    int first_child;  /* -1: leaf */
    int next;         /* threaded successor (-1 ends the walk) */
    int index;        /* first ion of a leaf (-1: empty) */
    int bucket_next;  /* unused for cells; see ion_next */
    int depth;
    float lx, ly, lz, dim;  /* location (corner) and side */
    float charge, uq;
    float cx, cy, cz;
    float delta;
} bh_cell;

typedef struct {
    bh_cell* c;
    int n, cap;
    int* ion_next;  /* bucket chains: next ion in the same depth-21 leaf */
} bh_tree;

static int bh_new_cell(bh_tree* t, float lx, float ly, float lz, float dim, int depth) {
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
    c->dim = dim;
    return t->n++;
}

/* locateSubcell: strict compares against location + dimensions / 2 */
static int bh_subcell(const bh_cell* c, const float* p) {
    const float h = c->dim / 2.0f;
    const float cx = c->lx + h, cy = c->ly + h, cz = c->lz + h;
    return (p[0] > cx) + ((p[1] > cy) << 1) + ((p[2] > cz) << 2);
}

/* generateSubcells + setLocationOfSubcells: 8 children, threaded */
static void bh_split(bh_tree* t, int ci) {
    const float h = t->c[ci].dim / 2.0f;
    int first = -1;
    for (int i = 0; i < 8; i++) {
        const bh_cell* p = &t->c[ci];
        float lx = p->lx, ly = p->ly, lz = p->lz;
        lx += (float)(i & 1) * h;
        ly += (float)((i & 2) >> 1) * h;
        lz += (float)((i & 4) >> 2) * h;
        const int k = bh_new_cell(t, lx, ly, lz, h, t->c[ci].depth + 1);
        if (i == 0) first = k;
    }
    for (int i = 0; i < 7; i++) t->c[first + i].next = first + i + 1;
    t->c[first + 7].next = t->c[ci].next;
    t->c[ci].first_child = first;
}

static void bh_add(bh_tree* t, const float* pos4, int ci, int ion) {
    for (;;) {
        if (t->c[ci].index == -1) {
            t->c[ci].index = ion;
            return;
        }
        if (t->c[ci].depth >= BH_MAX_DEPTH) {  /* bucket: append to the chain */
            int j = t->c[ci].index;
            while (t->ion_next[j] >= 0) j = t->ion_next[j];
            t->ion_next[j] = ion;
            return;
        }
        const int old = t->c[ci].index;
        bh_split(t, ci);
        t->c[ci].index = -1;
        const int s1 = bh_subcell(&t->c[ci], pos4 + 4 * (size_t)old);
        t->c[t->c[ci].first_child + s1].index = old;
        const int s2 = bh_subcell(&t->c[ci], pos4 + 4 * (size_t)ion);
        if (s1 != s2) {
            t->c[t->c[ci].first_child + s2].index = ion;
            return;
        }
        ci = t->c[ci].first_child + s1;  /* addToCell(temp, index) */
    }
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
            } else { /* bucket: same sums as a cell with the ions as children */
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
                    const float h = c->dim / 2.0f;
                    c->cx = c->lx + h;
                    c->cy = c->ly + h;
                    c->cz = c->lz + h;
                }
            }
        }
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
            const float h = c->dim / 2.0f;
            c->cx = c->lx + h;
            c->cy = c->ly + h;
            c->cz = c->lz + h;
        }
    }
    /* empty leaves keep charge 0; their centre is the geometric centre */
    if (c->first_child < 0 && c->index < 0) {
        const float h = c->dim / 2.0f;
        c->cx = c->lx + h;
        c->cy = c->ly + h;
        c->cz = c->lz + h;
    }
    const float h = c->dim / 2.0f;
    const float dx = c->cx - (c->lx + h), dy = c->cy - (c->ly + h), dz = c->cz - (c->lz + h);
    c->delta = sqrtf(dx * dx + dy * dy + dz * dz);
}

/* computeEffectiveDistance: float squares, double softening and sqrt */
static float bh_dist(const float* p, const bh_cell* c, float* rx, float* ry, float* rz) {
    *rx = p[0] - c->cx;
    *ry = p[1] - c->cy;
    *rz = p[2] - c->cz;
    const float s = *rx * *rx + *ry * *ry + *rz * *rz;
    return (float)sqrt((double)s + 6.4e-19);
}

static float bh_dist_pt(const float* p, const float* q, float* rx, float* ry, float* rz) {
    *rx = p[0] - q[0];
    *ry = p[1] - q[1];
    *rz = p[2] - q[2];
    const float s = *rx * *rx + *ry * *ry + *rz * *rz;
    return (float)sqrt((double)s + 6.4e-19);
}

/* One walk for target position p (already transformed); (cb, sb) rotate the
 * contributions back (1, 0 for none); ions equal to `self` are skipped. */
static void bh_walk(const bh_tree* t, const float* pos4, const float* p, int self, float theta,
                    int cb, int sb, float* acc) {
    int ci = 0;
    while (ci >= 0) {
        const bh_cell* c = &t->c[ci];
        if (c->first_child < 0) {
            if (c->index >= 0) {
                if (t->ion_next[c->index] < 0) {
                    if (c->index != self) {
                        float rx, ry, rz;
                        const float d = bh_dist(p, c, &rx, &ry, &rz);
                        const float f = c->charge / (d * d * d);
                        acc[0] += f * ((float)cb * rx - (float)sb * ry);
                        acc[1] += f * ((float)sb * rx + (float)cb * ry);
                        acc[2] += f * rz;
                    }
                } else {
                    for (int j = c->index; j >= 0; j = t->ion_next[j]) {
                        if (j == self) continue;
                        float rx, ry, rz;
                        const float* q = pos4 + 4 * (size_t)j;
                        const float d = bh_dist_pt(p, q, &rx, &ry, &rz);
                        const float f = q[3] / (d * d * d);
                        acc[0] += f * ((float)cb * rx - (float)sb * ry);
                        acc[1] += f * ((float)sb * rx + (float)cb * ry);
                        acc[2] += f * rz;
                    }
                }
            }
            ci = c->next;
        } else {
            float rx, ry, rz;
            const float d = bh_dist(p, c, &rx, &ry, &rz);
            if (theta * (d - c->delta) > c->dim) {
                const float f = c->charge / (d * d * d);
                acc[0] += f * ((float)cb * rx - (float)sb * ry);
                acc[1] += f * ((float)sb * rx + (float)cb * ry);
                acc[2] += f * rz;
                ci = c->next;
            } else {
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

/* Root cube shared with the GPU engine: bounding box of the sources in FP64,
 * side = largest extent x (1 + 2^-20) (1 if zero), centred on the box; the
 * float root is (centre - side/2, side) rounded to float. */
void orc_bh_root(const float* pos4, const float* acc4, int i0, int i1, float* root4) {
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    int any = 0;
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        any = 1;
        for (int c = 0; c < 3; c++) {
            const double v = (double)pos4[4 * (size_t)i + c];
            mn[c] = v < mn[c] ? v : mn[c];
            mx[c] = v > mx[c] ? v : mx[c];
        }
    }
    if (!any) {
        for (int c = 0; c < 3; c++) mn[c] = mx[c] = 0.0;
    }
    double side = 0.0;
    for (int c = 0; c < 3; c++) side = (mx[c] - mn[c]) > side ? (mx[c] - mn[c]) : side;
    side *= 1.0 + ldexp(1.0, -20);
    if (!(side > 0.0)) side = 1.0;
    for (int c = 0; c < 3; c++) root4[c] = (float)(0.5 * (mn[c] + mx[c]) - 0.5 * side);
    root4[3] = (float)side;
}

/*
 * Barnes-Hut accelerations of targets (status 1) in [i0, i1) from sources
 * (status 1 or 2) in the same range.  sym: 0 none, 1 rotational (four
 * quadrants), 2 translational (nimg lattice points times domain, in the
 * given order).  out3[3 (i - i0) + c] = the float acceleration the
 * reference accumulates (zero start), as double; zero for non-targets.
 * Returns the number of cells of the tree (including empty ones).
 */
int orc_bh_ref(const float* pos4, const float* vel4, const float* acc4, int n, int i0, int i1,
               float theta, int sym, int nimg, const int* lattice, const double* domain,
               double* out3) {
    (void)n;
    bh_tree t = {0};
    t.ion_next = (int*)malloc(sizeof(int) * (size_t)(i1 > 0 ? i1 : 1));
    for (int i = 0; i < i1; i++) t.ion_next[i] = -1;
    float root[4];
    orc_bh_root(pos4, acc4, i0, i1, root);
    bh_new_cell(&t, root[0], root[1], root[2], root[3], 0);
    for (int i = i0; i < i1; i++) {
        const float s = acc4[4 * (size_t)i + 3];
        if (!(s == 1.0f || s == 2.0f)) continue;
        int ci = 0;
        while (t.c[ci].first_child >= 0) ci = t.c[ci].first_child + bh_subcell(&t.c[ci], pos4 + 4 * (size_t)i);
        bh_add(&t, pos4, ci, i);
    }
    bh_props(&t, pos4, 0);
    for (int i = i0; i < i1; i++) {
        double* o = out3 + 3 * (size_t)(i - i0);
        o[0] = o[1] = o[2] = 0.0;
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
            bh_walk(&t, pos4, tp, self, theta, cb, sb, acc);
            for (int c = 0; c < 3; c++) tot[c] += (float)((double)acc[c] * kc);
        }
        for (int c = 0; c < 3; c++) o[c] = (double)tot[c];
    }
    const int ncell = t.n;
    free(t.c);
    free(t.ion_next);
    return ncell;
}
