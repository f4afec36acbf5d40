/*
 * Oracle (test infrastructure only): FP64 restatement of the reference's
 * energy bookkeeping, used to check the GPU engine.
 *
 * Follows particles.cpp:
 *   - computeChargedKineticEnergy :535-542 (0.5 amu m v^2)
 *   - computeChargedElectricPotentialEnergy :557-578 (sum over j != n of
 *     q_n q_j / sqrt(r^2 + eps^2), times e^2 / (4 pi eps0)), halved in
 *     Integrator::updateEnergyIntervalData (integration_schemes.cpp:799-835)
 *   - updateChargedFieldEnergy :598-638 (-e q times the left Riemann sums of
 *     E_z dz along z, then E_x dx + E_y dy along the radial leg, with
 *     numSteps points per leg)
 * in FP64 (the reference works in float); sums in long double.
 */
#include <math.h>

extern double orc_k_const(void);
extern double orc_species_mass(int s);
extern int orc_efield_f64(double x, double y, double z, double r0, double E0, const double* xy,
                          int ntri, const int* tri, const int* nbr, const double* er,
                          const double* ez, int* tri_io, double out[3]);

static const double ORC_E_CHARGE = (double)1.6021766e-19f;
static const double ORC_AMU_D = (double)1.66053904e-27f;

/* out[0] = kinetic energy, out[1] = potential energy (J) of the ions with
 * status 1 or 2. */
void orc_energy_f64(int n, const double* x, const double* y, const double* z, const double* q,
                    const double* vx, const double* vy, const double* vz, const int* status,
                    const int* species, double* out) {
    long double ke = 0.0L, pe = 0.0L;
    for (int i = 0; i < n; i++) {
        if (!(status[i] == 1 || status[i] == 2)) continue;
        const double v2 = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
        ke += (long double)(0.5 * ORC_AMU_D * orc_species_mass(species[i]) * v2);
        for (int j = i + 1; j < n; j++) {
            if (!(status[j] == 1 || status[j] == 2)) continue;
            const double dx = x[i] - x[j], dy = y[i] - y[j], dz = z[i] - z[j];
            pe += (long double)(q[i] * q[j] / sqrt(dx * dx + dy * dy + dz * dz + 6.4e-19));
        }
    }
    out[0] = (double)ke;
    out[1] = (double)(pe * (long double)(orc_k_const() * ORC_AMU_D));
}

/* fe[i] += -e q_i (sum_k E_z(x0, y0, z0 + k dz/ns) dz/ns
 *                + sum_k (E_x dx/ns + E_y dy/ns)(x0 + k dx/ns, y0 + k dy/ns, z)),
 * k = 0..ns-1, ns = nsteps_arr[i] (or nsteps). */
void orc_field_energy_f64(double r0, double E0, const double* xy, int ntri, const int* tri,
                          const int* nbr, const double* er, const double* ez, int n,
                          const double* x0, const double* y0, const double* z0, const double* x,
                          const double* y, const double* z, const double* q,
                          const int* nsteps_arr, int nsteps, double* fe) {
    for (int i = 0; i < n; i++) {
        const int ns = nsteps_arr ? nsteps_arr[i] : nsteps;
        const double dx = x[i] - x0[i], dy = y[i] - y0[i], dz = z[i] - z0[i];
        double fz = 0.0, fr = 0.0, e[3];
        int t = -1;
        for (int k = 0; k < ns; k++) {
            orc_efield_f64(x0[i], y0[i], z0[i] + k * dz / ns, r0, E0, xy, ntri, tri, nbr, er, ez,
                           &t, e);
            fz += e[2] * dz / ns;
        }
        const double zt = z0[i] + dz;
        for (int k = 0; k < ns; k++) {
            orc_efield_f64(x0[i] + k * dx / ns, y0[i] + k * dy / ns, zt, r0, E0, xy, ntri, tri,
                           nbr, er, ez, &t, e);
            fr += e[0] * dx / ns + e[1] * dy / ns;
        }
        fe[i] += (fz + fr) * (-ORC_E_CHARGE * q[i]);
    }
}
