/*
 * oracle/leapfrog.c -- CPU restatement of LeapfrogDKD on FP64 state.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle/direct.c).
 *
 * Follows LeapfrogDKD::runSimulation / driftNew / kick
 * (integration_schemes.cpp:705-751, :856-891), ChargedParticles::
 * propagatePositionNew / propagateVelocityNew (particles.cpp:693-786),
 * ParticleSystem::propagatePositionNew (particles.cpp:34-40: drift, symmetry
 * folding, boundary conditions), enforceSymmetryConditions (:199-340),
 * enforceBoundaryConditions (:73-177), computeBackgroundAcceleration
 * (:496-520) and findMostRecentInjected (:1328-1359, an ion takes part iff its
 * injection time is not after the drift's end time).  State is double; the
 * force is orc_direct_f64 and the field orc_efield_f64.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI_D 3.14159265358979323846

void orc_direct_f64(int n, const double* x, const double* y, const double* z, const double* q,
                    const int* status, const int* species, int i0, int i1, int sym, int nsec,
                    int nimg, const int* lattice, const double* domain, double* out3);
int orc_efield_f64(double x, double y, double z, double r0, double E0, const double* xy, int ntri,
                   const int* tri, const int* nbr, const double* er, const double* ez,
                   int* tri_io, double out[3]);
long double orc_species_mass_ld(int s);

typedef struct orc_lf_params {
    int sym, nsec;
    double dom[3];
    double bc[4];
    double bc_time;
    int use_efield;
    /* field mesh */
    const double* xy;
    const double* er;
    const double* ez;
    const int* tri;
    const int* nbr;
    int ntri;
    double r0, E0;
    /* translational images for the force */
    int nimg;
    const int* lattice;
} orc_lf_params;

static void fold_bound(const orc_lf_params* P, double* x, double* y, double* z, double* vx,
                       double* vy, int* st, double t_now, double t_inj) {
    if (P->sym == 1) {
        if (P->nsec == 4) {
            double tx, ty;
            if (*x < 0 && *y < 0) {
                *x = -*x; *y = -*y; *vx = -*vx; *vy = -*vy;
            } else if (*x < 0) {
                tx = *y; ty = -*x; *x = tx; *y = ty;
                tx = *vy; ty = -*vx; *vx = tx; *vy = ty;
            } else if (*y < 0) {
                tx = -*y; ty = *x; *x = tx; *y = ty;
                tx = -*vy; ty = *vx; *vx = tx; *vy = ty;
            }
        } else if (P->nsec > 1) {
            double w = 2.0 * ORC_PI_D / P->nsec;
            double phi = atan2(*y, *x);
            if (phi < 0) phi += 2.0 * ORC_PI_D;
            int k = (int)floor(phi / w);
            if (k > 0) {
                double c = cos(-k * w), s = sin(-k * w);
                double tx = c * *x - s * *y, ty = s * *x + c * *y;
                *x = tx; *y = ty;
                tx = c * *vx - s * *vy; ty = s * *vx + c * *vy;
                *vx = tx; *vy = ty;
            }
        }
    } else if (P->sym == 2) {
        double* p[3] = {x, y, z};
        for (int c = 0; c < 3; c++) {
            double D = P->dom[c];
            if (!(D > 0)) continue;
            double t = *p[c] + D / 2;
            if (t >= 0) {
                *p[c] = fmod(t, D) - D / 2;
            } else {
                t = fmod(t, D);
                *p[c] = t < 0 ? t + D / 2 : t - D / 2;
            }
        }
    }
    if (*st == 1) {
        int out = 0;
        if (P->bc[0] != 0 && fabs(*x) > P->bc[0]) out = 1;
        if (P->bc[1] != 0 && fabs(*y) > P->bc[1]) out = 1;
        if (P->bc[2] != 0 && fabs(*z) > P->bc[2]) out = 1;
        if (P->bc[3] != 0 && sqrt(*x * *x + *y * *y) > P->bc[3]) out = 1;
        if (P->bc_time != 0 && t_now - t_inj > P->bc_time) out = 1;
        if (out) *st = 0;
    }
}

/* driftNew(dt/2): propagate positions of injected ions, then fold and bound. */
void orc_drift_f64(const orc_lf_params* P, int n, double* x, double* y, double* z, double* vx,
                   double* vy, const double* vz, int* status, const double* tinj, double t_start,
                   double h) {
    double t_end = t_start + h;
    for (int i = 0; i < n; i++) {
        if (tinj[i] > t_end) continue;
        if (status[i] == 1 || status[i] == 2) {
            double s = fmin(fmax(t_end - tinj[i], 0.0), h);
            x[i] += vx[i] * s;
            y[i] += vy[i] * s;
            z[i] += vz[i] * s;
        }
        fold_bound(P, &x[i], &y[i], &z[i], &vx[i], &vy[i], &status[i], t_end, tinj[i]);
    }
}

/* kick(dt): Coulomb acceleration (given) plus the background field, then
 * propagateVelocityNew over the ions active since the drift to t + dt/2. */
void orc_kick_f64(const orc_lf_params* P, int n, const double* x, const double* y, const double* z,
                  double* vx, double* vy, double* vz, const double* q, const int* status,
                  const int* species, const double* tinj, const double* acc3, double t, double dt) {
    const double f0 = (double)1.6021766e-19f / (double)1.66053904e-27f;
    for (int i = 0; i < n; i++) {
        if (status[i] != 1 || tinj[i] > t + 0.5 * dt) continue;
        double ax = acc3[3 * i], ay = acc3[3 * i + 1], az = acc3[3 * i + 2];
        if (P->use_efield) {
            double e[3];
            int tri = -1;
            orc_efield_f64(x[i], y[i], z[i], P->r0, P->E0, P->xy, P->ntri, P->tri, P->nbr, P->er,
                           P->ez, &tri, e);
            double f = f0 * q[i] / (double)orc_species_mass_ld(species[i]);
            ax += f * e[0];
            ay += f * e[1];
            az += f * e[2];
        }
        double s = fmin(fmax(t + dt - tinj[i], 0.0), dt);
        vx[i] += ax * s;
        vy[i] += ay * s;
        vz[i] += az * s;
    }
}

/* nsteps of DKD with the FP64 direct force; acc3 receives the last force. */
void orc_leapfrog_run_f64(const orc_lf_params* P, int n, double* x, double* y, double* z,
                          double* vx, double* vy, double* vz, const double* q, int* status,
                          const int* species, const double* tinj, double t0, double dt, int nsteps,
                          double* acc3) {
    int* eff = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int k = 0; k < nsteps; k++) {
        double t = t0 + k * dt;
        orc_drift_f64(P, n, x, y, z, vx, vy, vz, status, tinj, t, 0.5 * dt);
        for (int i = 0; i < n; i++) eff[i] = tinj[i] <= t + 0.5 * dt ? status[i] : 0;
        orc_direct_f64(n, x, y, z, q, eff, species, 0, n, P->sym, P->nsec, P->nimg, P->lattice,
                       P->dom, acc3);
        orc_kick_f64(P, n, x, y, z, vx, vy, vz, q, status, species, tinj, acc3, t, dt);
        orc_drift_f64(P, n, x, y, z, vx, vy, vz, status, tinj, t + 0.5 * dt, 0.5 * dt);
    }
    free(eff);
}
