/*
 * Oracle (test infrastructure only): FP64 derivatives of the direct force,
 * used to check the GPU engine's force JVP/VJP.  The reference has no
 * derivatives; this differentiates its FP64-restated direct sum
 * (direct_force_method.cpp:153-190):
 *   a_i = k_i sum_{j != i} q_j G(r_i - r_j),  k_i = K q_i / m_i,
 *   G(r) = r s^-3/2,  H(r) = I s^-3/2 - 3 r r^T s^-5/2,  s = |r|^2 + eps^2.
 * Targets have status 1, sources status 1 or 2.
 */
#include <math.h>
#include <stddef.h>

extern double orc_k_const(void);
extern double orc_species_mass(int s);

static int src_ok(int s) { return s == 1 || s == 2; }

/* JVP: tangents (dx, dy, dz, dq) -> tangent of a (zero for non-targets). */
void orc_force_jvp_f64(int n, const double* x, const double* y, const double* z, const double* q,
                       const int* st, const int* sp, const double* dx, const double* dy,
                       const double* dz, const double* dq, double* out3) {
    const double K = orc_k_const();
    for (int i = 0; i < n; i++) {
        double* o = out3 + 3 * (size_t)i;
        o[0] = o[1] = o[2] = 0.0;
        if (st[i] != 1) continue;
        double h[3] = {0, 0, 0}, g[3] = {0, 0, 0}, c[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (j == i || !src_ok(st[j])) continue;
            const double r[3] = {x[i] - x[j], y[i] - y[j], z[i] - z[j]};
            const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + 6.4e-19;
            const double s32 = pow(s, -1.5), s52 = pow(s, -2.5);
            const double e[3] = {dx[i] - dx[j], dy[i] - dy[j], dz[i] - dz[j]};
            const double re = r[0] * e[0] + r[1] * e[1] + r[2] * e[2];
            for (int a = 0; a < 3; a++) {
                h[a] += q[j] * (e[a] * s32 - 3.0 * r[a] * re * s52);
                g[a] += dq[j] * r[a] * s32;
                c[a] += q[j] * r[a] * s32;
            }
        }
        const double m = orc_species_mass(sp[i]);
        const double ki = K * q[i] / m, kd = K * dq[i] / m;
        for (int a = 0; a < 3; a++) o[a] = ki * (h[a] + g[a]) + kd * c[a];
    }
}

/* VJP: cotangent b (3 per ion) of a -> cotangents of positions (3 per ion)
 * and charges. */
void orc_force_vjp_f64(int n, const double* x, const double* y, const double* z, const double* q,
                       const int* st, const int* sp, const double* b3, double* xb3, double* qb) {
    const double K = orc_k_const();
    for (int k = 0; k < n; k++) {
        double* o = xb3 + 3 * (size_t)k;
        o[0] = o[1] = o[2] = 0.0;
        qb[k] = 0.0;
        if (!src_ok(st[k])) continue;
        double uk[3] = {0, 0, 0};
        if (st[k] == 1) {
            const double kk = K * q[k] / orc_species_mass(sp[k]);
            for (int a = 0; a < 3; a++) uk[a] = kk * b3[3 * (size_t)k + a];
        }
        double gsum = 0.0, c[3] = {0, 0, 0};
        for (int j = 0; j < n; j++) {
            if (j == k || !src_ok(st[j])) continue;
            double uj[3] = {0, 0, 0};
            if (st[j] == 1) {
                const double kj = K * q[j] / orc_species_mass(sp[j]);
                for (int a = 0; a < 3; a++) uj[a] = kj * b3[3 * (size_t)j + a];
            }
            const double r[3] = {x[k] - x[j], y[k] - y[j], z[k] - z[j]};
            const double s = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] + 6.4e-19;
            const double s32 = pow(s, -1.5), s52 = pow(s, -2.5);
            /* H (q_j u_k - q_k u_j) */
            double w[3];
            for (int a = 0; a < 3; a++) w[a] = q[j] * uk[a] - q[k] * uj[a];
            const double rw = r[0] * w[0] + r[1] * w[1] + r[2] * w[2];
            for (int a = 0; a < 3; a++) o[a] += w[a] * s32 - 3.0 * r[a] * rw * s52;
            gsum += (uj[0] * r[0] + uj[1] * r[1] + uj[2] * r[2]) * s32;
            for (int a = 0; a < 3; a++) c[a] += q[j] * r[a] * s32;
        }
        qb[k] = -gsum;
        if (st[k] == 1) {
            const double km = K / orc_species_mass(sp[k]);
            qb[k] += km * (b3[3 * (size_t)k] * c[0] + b3[3 * (size_t)k + 1] * c[1] +
                           b3[3 * (size_t)k + 2] * c[2]);
        }
    }
}
