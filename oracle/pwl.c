/*
 * oracle/pwl.c -- CPU restatement of the reference's background E-field
 * (piecewise-linear interpolation on a scattered (r, z) triangulation).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/direct.c).
 *
 * Follows:
 *   triangulation_search_delaunay   pwl_interp_2d_scattered.c:1762-1990
 *     (barycentric walk; neighbour k lies across the edge opposite vertex k,
 *      negative = hull; the reference keeps a static warm-start triangle, which
 *      here is an explicit in/out argument so call history is reproducible)
 *   pwl_interp_2d_scattered_value   pwl_interp_2d_scattered.c:887-956
 *   computeEfieldSingleEmitter      interpolation_scheme.cpp:200-271
 *   computeBackgroundAcceleration   particles.cpp:496-520
 * Compile with -ffp-contract=off.
 */
#include <math.h>
#include <stdio.h>

/* Walk from *tri (or from (ntri+1)/2-1 when *tri is out of range) toward p.
 * Returns the edge code (0 inside or on the boundary, -1..-3 outside the hull
 * past that edge, -4 on a dead end, -5 when the step limit is hit) and leaves
 * the final triangle in *tri and the barycentric coordinates in bc[3]. */
int orc_tri_search(const double* xy, int ntri, const int* tri, const int* nbr,
                   const double p[2], int* tri_io, double bc[3], int* steps) {
    int t = *tri_io;
    if (t < 0 || ntri <= t) t = (ntri + 1) / 2 - 1;
    int step = -1;
    int edge = 0;
    for (;;) {
        step++;
        if (ntri < step) { edge = -5; break; }
        int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
        double dxa = xy[2 * a] - xy[2 * c];
        double dya = xy[2 * a + 1] - xy[2 * c + 1];
        double dxb = xy[2 * b] - xy[2 * c];
        double dyb = xy[2 * b + 1] - xy[2 * c + 1];
        double dxp = p[0] - xy[2 * c];
        double dyp = p[1] - xy[2 * c + 1];
        double det = dxa * dyb - dya * dxb;
        double al = (dxp * dyb - dyp * dxb) / det;
        double be = (dxa * dyp - dya * dxp) / det;
        double ga = 1.0 - al - be;
        bc[0] = al; bc[1] = be; bc[2] = ga;
        if (0.0 <= al && 0.0 <= be && 0.0 <= ga) break;
        if (al < 0.0 && 0 <= nbr[3 * t]) { t = nbr[3 * t]; continue; }
        else if (be < 0.0 && 0 <= nbr[3 * t + 1]) { t = nbr[3 * t + 1]; continue; }
        else if (ga < 0.0 && 0 <= nbr[3 * t + 2]) { t = nbr[3 * t + 2]; continue; }
        if (al < 0.0) { edge = -2; break; }
        else if (be < 0.0) { edge = -3; break; }
        else if (ga < 0.0) { edge = -1; break; }
        edge = -4;
        break;
    }
    *tri_io = t;
    if (steps) *steps = step;
    return edge;
}

/* pwl value at p for two value arrays at once (the reference walks twice; the
 * second walk starts inside the first one's triangle and stops there). */
int orc_pwl_value2(const double* xy, int ntri, const int* tri, const int* nbr,
                   const double* z1, const double* z2, const double p[2], int* tri_io,
                   double* v1, double* v2) {
    double bc[3];
    int edge = orc_tri_search(xy, ntri, tri, nbr, p, tri_io, bc, 0);
    int t = *tri_io;
    int a = tri[3 * t], b = tri[3 * t + 1], c = tri[3 * t + 2];
    *v1 = bc[0] * z1[a] + bc[1] * z1[b] + bc[2] * z1[c];
    *v2 = bc[0] * z2[a] + bc[1] * z2[b] + bc[2] * z2[c];
    return edge;
}

/* Reference contract of computeEfieldSingleEmitter on float inputs.  r0/E0 are
 * the float members of InterpolationScheme.  out[3] float. */
void orc_efield_ref(float x, float y, float z, float r0, float E0, const double* xy, int ntri,
                    const int* tri, const int* nbr, const double* er, const double* ez,
                    int* tri_io, float out[3]) {
    const long double PI = 3.14159265358979323846264338328L;
    float phi = atanf(y / x);
    if (x < 0) phi = (float)((long double)phi + PI);
    if (x == 0 && y != 0) phi = (float)(PI / 2);
    else if (x == 0 && y == 0) phi = 0;
    x /= r0;
    y /= r0;
    z /= r0;
    double coords[2];
    coords[0] = sqrt((double)(x * x + y * y));
    coords[1] = (double)z;
    float e[3];
    if (z < 0 || z > 140) {
        e[0] = e[1] = e[2] = 0.0f;
    } else {
        double vr, vz;
        orc_pwl_value2(xy, ntri, tri, nbr, er, ez, coords, tri_io, &vr, &vz);
        e[0] = (float)vr;
        e[0] *= cosf(phi);
        e[1] = (float)vr;
        e[1] *= sinf(phi);
        e[2] = (float)vz;
    }
    out[0] = e[0] * E0;
    out[1] = e[1] * E0;
    out[2] = e[2] * E0;
}

/* computeBackgroundAcceleration (particles.cpp:496-520) over [i0, i1) in index
 * order with the warm start chained from call to call, like the reference.
 * Adds into acc4 xyz (float) and writes the field into efield3 (float, may be 0). */
void orc_background_accel_ref(const float* pos4, const float* vel4, float* acc4, int i0, int i1,
                              float r0, float E0, const double* xy, int ntri, const int* tri,
                              const int* nbr, const double* er, const double* ez, int* tri_io,
                              float* efield3) {
    const double factor = (double)1.6021766e-19f / (double)1.66053904e-27f;
    extern long double orc_species_mass_ld(int s);
    for (int i = i0; i < i1; i++) {
        if (acc4[4 * i + 3] != 1.0f) continue;
        int sp = (int)vel4[4 * i + 3];
        long double m = sp == 3 ? 0.0L : orc_species_mass_ld(sp);
        float ef = (float)((long double)(factor * (double)pos4[4 * i + 3]) / m);
        float e[3];
        orc_efield_ref(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], r0, E0, xy, ntri, tri, nbr,
                       er, ez, tri_io, e);
        if (efield3) { efield3[3 * i] = e[0]; efield3[3 * i + 1] = e[1]; efield3[3 * i + 2] = e[2]; }
        e[0] *= ef; e[1] *= ef; e[2] *= ef;
        acc4[4 * i] += e[0];
        acc4[4 * i + 1] += e[1];
        acc4[4 * i + 2] += e[2];
    }
}

/* FP64-state field: the same interpolant in double.  cos/sin of the azimuth are
 * x/rho and y/rho (phi = 0 on the axis), the gate is z/r0 in [0, 140]. */
int orc_efield_f64(double x, double y, double z, double r0, double E0, const double* xy, int ntri,
                   const int* tri, const int* nbr, const double* er, const double* ez,
                   int* tri_io, double out[3]) {
    double rho = sqrt(x * x + y * y);
    double zn = z / r0;
    out[0] = out[1] = out[2] = 0.0;
    if (zn < 0 || zn > 140) return 0;
    double p[2] = {rho / r0, zn};
    double vr, vz;
    int edge = orc_pwl_value2(xy, ntri, tri, nbr, er, ez, p, tri_io, &vr, &vz);
    double c = 1.0, s = 0.0;
    if (rho > 0) { c = x / rho; s = y / rho; }
    out[0] = E0 * vr * c;
    out[1] = E0 * vr * s;
    out[2] = E0 * vz;
    return edge;
}

/* Batch of orc_efield_f64, each query walking from the caller's start triangle
 * seeds[i] (or the default start when seeds is 0); returns edge codes. */
void orc_efield_f64_batch(int n, const double* x, const double* y, const double* z, double r0,
                          double E0, const double* xy, int ntri, const int* tri, const int* nbr,
                          const double* er, const double* ez, const int* seeds, double* out3,
                          int* edges, int* tris) {
    for (int i = 0; i < n; i++) {
        int t = seeds ? seeds[i] : -1;
        int e = orc_efield_f64(x[i], y[i], z[i], r0, E0, xy, ntri, tri, nbr, er, ez, &t, out3 + 3 * i);
        if (edges) edges[i] = e;
        if (tris) tris[i] = t;
    }
}
