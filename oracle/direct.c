/*
 * oracle/direct.c -- CPU restatement of the reference's softened Coulomb sum.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP force engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (libsolveit_hip.so) never links or calls it.
 *
 * Two arithmetic contracts are restated:
 *
 *  (1) "ref" -- the reference's own arithmetic on its float AoS state
 *      (DirectForceFunction::direct, direct_force_method.cpp:153-190):
 *        dist   = (float)(x_i - x_j)            float - float, then widened
 *        invD   = 1.0 / sqrt(|dist|^2 + eps)     double
 *        invD3  = q_j * invD * invD * invD       double, left to right
 *        a_i   += dist * invD3                   double, j ascending
 *        a_i   *= K * q_i / m_i                  long double division -> double
 *      Rotational sectors follow direct_force_method_rotational.cpp:48-94
 *      (source rotated in float by integer cos/sin for the 4-fold case; no
 *      j != i exclusion), translational images follow
 *      direct_force_method_translational.cpp:320-367 (target shifted by the
 *      float lattice offset in float first).  The reference rounds each
 *      image's/sector's scaled double sum into the float acceleration
 *      (acceleration[i].x += ai.x); orc_direct_ref() returns the
 *      pre-rounding double sum over images, and orc_direct_ref_accumulate()
 *      additionally replays the float accumulation in the caller's image order.
 *
 *  (2) "f64" -- the same physics on double positions with double displacements
 *      (the accuracy oracle for the engine's FP64-state paths).
 *
 * Compile with -ffp-contract=off: the reference is built for x86-64 without FMA,
 * so every product and sum above is rounded separately.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define ORC_STATUS_INBOUND 1.0f   /* structures.h:206-211 */
#define ORC_STATUS_AFFECTS 2.0f

/* constants.h:28-47.  ELEM_CHARGE*ELEM_CHARGE is a float product; the
 * denominator 4*PI*EPS0*AMU is evaluated in long double (PI is long double). */
static const long double ORC_PI = 3.14159265358979323846264338328L;
static const float ORC_ELEM_CHARGE = 1.6021766e-19f;
static const float ORC_AMU = 1.66053904e-27f;
static const float ORC_EPS0 = 8.85419e-12f;
static const double ORC_SOFTENING = 6.4e-19;

double orc_k_const(void) {
    float e2 = ORC_ELEM_CHARGE * ORC_ELEM_CHARGE;
    long double den = 4 * ORC_PI * ORC_EPS0 * ORC_AMU;
    return (double)((long double)e2 / den);
}

double orc_softening(void) { return ORC_SOFTENING; }

/* ELEM_CHARGE / AMU in double (particles.cpp:500). */
double orc_e_over_amu(void) { return (double)ORC_ELEM_CHARGE / (double)ORC_AMU; }

/* ParticleSpeciesFunction::getMass, particles.h:67-87 (amu). */
long double orc_species_mass_ld(int s) {
    const double EMI = 111, BF4 = 87;
    switch (s) {
    case 0: return EMI;
    case 1: return 2 * EMI + BF4;
    case 2: return 3 * EMI + 2 * BF4;
    case 3: return -1;
    case 4: return EMI + BF4;
    case 5: return BF4;
    case 6: return EMI + 2 * BF4;
    case 7: return 2 * EMI + 3 * BF4;
    }
    return 0;
}
double orc_species_mass(int s) { return (double)orc_species_mass_ld(s); }

/* accel_const = K_const * q_i / getMass(i): double product, long double
 * division (getMass returns long double), rounded to double. */
static double accel_const_ref(float q, int species) {
    double kq = orc_k_const() * (double)q;
    long double m = species == 3 ? 0.0L : orc_species_mass_ld(species);
    return (double)((long double)kq / m);
}

/* quadrantRotationFunction, auxiliary_functions.h:55-57 */
static int qrf(int n) {
    int oe = (n & 1) ? -1 : 1;
    int a = 1 - oe;
    int sg = (a > 0) - (a < 0);
    int oe2 = (((n - 1) >> 1) & 1) ? -1 : 1;
    return sg * oe2;
}
int orc_quadrant_cos(int k) { return qrf(k + 1); }
int orc_quadrant_sin(int k) { return qrf(k); }

/* Sector rotation constants for k of nsec sectors, in float (ref contract).
 * For nsec == 4 they are the reference's exact integers. */
static void sector_cs_f(int k, int nsec, float* c, float* s) {
    if (nsec == 4) { *c = (float)qrf(k + 1); *s = (float)qrf(k); return; }
    double ang = 2.0 * (double)ORC_PI * k / nsec;
    *c = (float)cos(ang); *s = (float)sin(ang);
}
static void sector_cs_d(int k, int nsec, double* c, double* s) {
    if ((4 * k) % nsec == 0) {           /* exact multiples of pi/2 */
        int q = (4 * k) / nsec;
        *c = (double)qrf(q + 1); *s = (double)qrf(q); return;
    }
    double ang = 2.0 * (double)ORC_PI * k / nsec;
    *c = cos(ang); *s = sin(ang);
}

/* One (image, target) double sum in the ref contract. sym: 0 plain, 1 rotational
 * sector (c,s), 2 translational (off).  Returns the unscaled sum in ai[3]. */
static void ref_target_sum(int i, const float* pos4, const float* acc4, int j0, int j1,
                           int sym, float c, float s, const float off[3], double ai[3]) {
    ai[0] = ai[1] = ai[2] = 0.0;
    const float xi = pos4[4 * i], yi = pos4[4 * i + 1], zi = pos4[4 * i + 2];
    for (int j = j0; j < j1; j++) {
        float st = acc4[4 * j + 3];
        if (!(st == ORC_STATUS_INBOUND || st == ORC_STATUS_AFFECTS)) continue;
        const float xj = pos4[4 * j], yj = pos4[4 * j + 1], zj = pos4[4 * j + 2];
        double dx, dy, dz;
        if (sym == 0) {
            if (j == i) continue;
            dx = (double)(float)(xi - xj);
            dy = (double)(float)(yi - yj);
            dz = (double)(float)(zi - zj);
        } else if (sym == 1) {
            float rx = (float)(c * xj) - (float)(s * yj);
            float ry = (float)(s * xj) + (float)(c * yj);
            dx = (double)(float)(xi - rx);
            dy = (double)(float)(yi - ry);
            dz = (double)(float)(zi - zj);
        } else {
            dx = (double)(float)((float)(xi - off[0]) - xj);
            dy = (double)(float)((float)(yi - off[1]) - yj);
            dz = (double)(float)((float)(zi - off[2]) - zj);
        }
        double invD = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + ORC_SOFTENING);
        double invD3 = (double)pos4[4 * j + 3] * invD * invD * invD;
        ai[0] += dx * invD3;
        ai[1] += dy * invD3;
        ai[2] += dz * invD3;
    }
}

/*
 * Reference-contract direct sum.  pos4/vel4/acc4 are the reference's AoS
 * vec4<float> arrays (position xyz+q, velocity xyz+species, acceleration
 * xyz+status).  Targets [i0,i1) with status inbound; sources [j0,j1) with status
 * inbound or affectsAcceleration.  sym: 0 plain, 1 rotational (nsec sectors),
 * 2 translational (nimg lattice points, domain = float cell size).
 * out3[3*(i-i0)+c] = pre-rounding double acceleration (0 for non-targets).
 */
void orc_direct_ref(const float* pos4, const float* vel4, const float* acc4,
                    int i0, int i1, int j0, int j1, int sym, int nsec,
                    int nimg, const int* lattice, const float* domain, double* out3) {
    int nim = sym == 0 ? 1 : (sym == 1 ? nsec : nimg);
    for (int i = i0; i < i1; i++) {
        double* o = out3 + 3 * (size_t)(i - i0);
        o[0] = o[1] = o[2] = 0.0;
        if (acc4[4 * i + 3] != ORC_STATUS_INBOUND) continue;
        double kc = accel_const_ref(pos4[4 * i + 3], (int)vel4[4 * i + 3]);
        for (int k = 0; k < nim; k++) {
            float c = 1.0f, s = 0.0f, off[3] = {0.0f, 0.0f, 0.0f};
            if (sym == 1) sector_cs_f(k, nsec, &c, &s);
            if (sym == 2) {
                off[0] = (float)lattice[3 * k] * domain[0];
                off[1] = (float)lattice[3 * k + 1] * domain[1];
                off[2] = (float)lattice[3 * k + 2] * domain[2];
            }
            double ai[3];
            ref_target_sum(i, pos4, acc4, j0, j1, sym, c, s, off, ai);
            o[0] += ai[0] * kc;
            o[1] += ai[1] * kc;
            o[2] += ai[2] * kc;
        }
    }
}

/* Replays the reference's float accumulation (acceleration[i].x += ai.x once
 * per image/sector, images in the caller's order) into acc4's xyz. */
void orc_direct_ref_accumulate(const float* pos4, const float* vel4, float* acc4,
                               int i0, int i1, int j0, int j1, int sym, int nsec,
                               int nimg, const int* lattice, const float* domain) {
    int nim = sym == 0 ? 1 : (sym == 1 ? nsec : nimg);
    for (int k = 0; k < nim; k++) {
        float c = 1.0f, s = 0.0f, off[3] = {0.0f, 0.0f, 0.0f};
        if (sym == 1) sector_cs_f(k, nsec, &c, &s);
        if (sym == 2) {
            off[0] = (float)lattice[3 * k] * domain[0];
            off[1] = (float)lattice[3 * k + 1] * domain[1];
            off[2] = (float)lattice[3 * k + 2] * domain[2];
        }
        for (int i = i0; i < i1; i++) {
            if (acc4[4 * i + 3] != ORC_STATUS_INBOUND) continue;
            double ai[3];
            ref_target_sum(i, pos4, acc4, j0, j1, sym, c, s, off, ai);
            double kc = accel_const_ref(pos4[4 * i + 3], (int)vel4[4 * i + 3]);
            ai[0] *= kc; ai[1] *= kc; ai[2] *= kc;
            acc4[4 * i] = (float)((double)acc4[4 * i] + ai[0]);
            acc4[4 * i + 1] = (float)((double)acc4[4 * i + 1] + ai[1]);
            acc4[4 * i + 2] = (float)((double)acc4[4 * i + 2] + ai[2]);
        }
    }
}

/*
 * FP64-state direct sum (double positions, double displacements).  Same
 * selection rules as the ref contract; status/species are int arrays; q in e.
 * sym 1 rotates sources about z by 2*pi*k/nsec; sym 2 shifts the target by
 * lattice*domain (double).  Sums run j ascending per image; images in order.
 */
void orc_direct_f64(int n, const double* x, const double* y, const double* z,
                    const double* q, const int* status, const int* species,
                    int i0, int i1, int sym, int nsec, int nimg, const int* lattice,
                    const double* domain, double* out3) {
    int nim = sym == 0 ? 1 : (sym == 1 ? nsec : nimg);
    for (int i = i0; i < i1; i++) {
        double* o = out3 + 3 * (size_t)(i - i0);
        o[0] = o[1] = o[2] = 0.0;
        if (status[i] != 1) continue;
        double kc = (double)((long double)(orc_k_const() * q[i]) /
                             (species[i] == 3 ? 0.0L : orc_species_mass_ld(species[i])));
        for (int k = 0; k < nim; k++) {
            double c = 1.0, s = 0.0, ox = 0.0, oy = 0.0, oz = 0.0;
            if (sym == 1) sector_cs_d(k, nsec, &c, &s);
            if (sym == 2) {
                ox = lattice[3 * k] * domain[0];
                oy = lattice[3 * k + 1] * domain[1];
                oz = lattice[3 * k + 2] * domain[2];
            }
            double ax = 0, ay = 0, az = 0;
            const double xi = x[i] - ox, yi = y[i] - oy, zi = z[i] - oz;
            for (int j = 0; j < n; j++) {
                if (!(status[j] == 1 || status[j] == 2)) continue;
                if (sym == 0 && j == i) continue;
                double xj = x[j], yj = y[j];
                if (sym == 1) { xj = c * x[j] - s * y[j]; yj = s * x[j] + c * y[j]; }
                double dx = xi - xj, dy = yi - yj, dz = zi - z[j];
                double invD = 1.0 / sqrt(dx * dx + dy * dy + dz * dz + ORC_SOFTENING);
                double invD3 = q[j] * invD * invD * invD;
                ax += dx * invD3; ay += dy * invD3; az += dz * invD3;
            }
            o[0] += ax * kc; o[1] += ay * kc; o[2] += az * kc;
        }
    }
}

/* Same as orc_direct_f64 plain, but with every product, sum and the square root
 * in long double: the brute-force check used to pin the double oracle. */
void orc_direct_f64_ld(int n, const double* x, const double* y, const double* z,
                       const double* q, const int* status, const int* species,
                       double* out3) {
    long double K = (long double)orc_k_const();
    for (int i = 0; i < n; i++) {
        long double ax = 0, ay = 0, az = 0;
        out3[3 * i] = out3[3 * i + 1] = out3[3 * i + 2] = 0.0;
        if (status[i] != 1) continue;
        for (int j = 0; j < n; j++) {
            if (j == i || !(status[j] == 1 || status[j] == 2)) continue;
            long double dx = (long double)x[i] - x[j], dy = (long double)y[i] - y[j],
                        dz = (long double)z[i] - z[j];
            long double r2 = dx * dx + dy * dy + dz * dz + (long double)ORC_SOFTENING;
            long double inv = 1.0L / sqrtl(r2);
            long double w = (long double)q[j] * inv * inv * inv;
            ax += dx * w; ay += dy * w; az += dz * w;
        }
        long double kc = K * (long double)q[i] / orc_species_mass_ld(species[i]);
        out3[3 * i] = (double)(ax * kc);
        out3[3 * i + 1] = (double)(ay * kc);
        out3[3 * i + 2] = (double)(az * kc);
    }
}
