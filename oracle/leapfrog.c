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

/*
 * LeapfrogDKDSubsteps (integration_schemes.cpp:899-1529) on FP64 state.
 *
 * Levels 0..M with dt_l = dt_{l-1} / 3 (setSubstepTimes :1375-1413); each ion
 * carries a level (particleLevels, initially M).  One iteration is the
 * recursion runSimulation :1497-1524 -> driftNew(0) kick(0, 0) driftNew(0),
 * with, for L < M, driftNew(L) = driftNew(L+1) kick(L+1, L+1) driftNew(L+1)
 * [crossoverParticles(L) when it is the second drift of its parent] and
 * kick(L, T) = driftNew(L+1) kick(L+1, T) driftNew(L+1) (:1135-1173, :1252-1296);
 * at L = M a drift moves every injected ion by dt_M / 2 (propagatePositionNew)
 * and a kick sets the status of the injected ions from their levels
 * (setParticleStatus :1193-1211: level >= T -> 1, below T -> 2; out and
 * crashed ions untouched), evaluates the force on that status (the main
 * method when T = 0, the substep method otherwise), adds the background
 * field and kicks each status-1 ion by its own level's dt, clamped from the
 * injection time with the window ending at tempVelTime + dt_level
 * (variableTimestepKick :1216-1250); tempVelTime then advances by dt_M.
 * crossoverParticles(L) (:1117-1127) re-levels the injected ions at level L
 * or L+1 by the boundary conditions of level L+1's type (checkConditions
 * :919-937, checkSpatialConditions :944-1066, checkTemporalConditions
 * :1073-1110).
 */
typedef struct orc_substep_level {
    int bc_type;         /* 0 none, 1 space, 2 time, 3 space and time */
    double bc_space[4];  /* |x|, |y|, |z|, rho limits (0: none) */
    double bc_time;      /* limit on the time since injection (0: none) */
} orc_substep_level;

/* The finest level (from M down to `level`) whose limit admits v, extended
 * down over the levels that share that limit; -1 when none admits it. */
static int substep_scan(const orc_substep_level* T, int M, int level, int axis, double v) {
    int possible = -1, found = 0;
    double bound = 0.0;
    for (int i = M; i >= level && !found; i--) {
        const double bc = axis < 4 ? T[i].bc_space[axis] : T[i].bc_time;
        if (possible == -1) {
            if (v < bc || bc == 0) {
                possible = i;
                bound = bc;
            }
        } else {
            if (bound == bc) possible--;
            else found = 1;
        }
    }
    return possible;
}

/* checkConditions(index, level): the new level of one ion. */
int orc_substep_level_of(const orc_substep_level* T, int M, int level, int cur, double x, double y,
                         double z, double tnow, double tinj) {
    const int type = T[level + 1].bc_type;
    if (type == 0) return cur;
    int sp = cur;
    if (type == 1 || type == 3) {
        const double v[4] = {fabs(x), fabs(y), fabs(z), sqrt(x * x + y * y)};
        sp = substep_scan(T, M, level, 0, v[0]);
        for (int a = 1; a < 4; a++) {
            const int p = substep_scan(T, M, level, a, v[a]);
            sp = p > sp ? p : sp;
        }
        if (type == 1) return sp;
    }
    const int tp = substep_scan(T, M, level, 4, tnow - tinj);
    if (type == 2) return tp;
    return tp > sp ? tp : sp;
}

typedef struct {
    const orc_lf_params* P;
    const orc_substep_level* T;
    int M, n;
    double *x, *y, *z, *vx, *vy, *vz;
    const double *q, *tinj;
    int* status;
    const int* species;
    int* level;
    double* dtl;
    double tpos, tvel;
    int* eff;
    double* acc3;
} substep_run;

static void ss_drift(substep_run* R) {
    const double h = R->dtl[R->M] / 2;
    orc_drift_f64(R->P, R->n, R->x, R->y, R->z, R->vx, R->vy, R->vz, R->status, R->tinj, R->tpos, h);
    R->tpos += h;
}

static void ss_kick(substep_run* R, int top) {
    const int n = R->n;
    for (int i = 0; i < n; i++) {
        if (R->tinj[i] > R->tpos) continue;  /* not injected yet */
        if (R->status[i] == 0 || R->status[i] == -1) continue;
        R->status[i] = R->level[i] >= top ? 1 : 2;
    }
    for (int i = 0; i < n; i++) R->eff[i] = R->tinj[i] <= R->tpos ? R->status[i] : 0;
    orc_direct_f64(n, R->x, R->y, R->z, R->q, R->eff, R->species, 0, n, R->P->sym, R->P->nsec,
                   R->P->nimg, R->P->lattice, R->P->dom, R->acc3);
    const double f0 = (double)1.6021766e-19f / (double)1.66053904e-27f;
    for (int i = 0; i < n; i++) {
        if (R->eff[i] != 1) continue;
        double ax = R->acc3[3 * i], ay = R->acc3[3 * i + 1], az = R->acc3[3 * i + 2];
        if (R->P->use_efield) {
            double e[3];
            int tri = -1;
            orc_efield_f64(R->x[i], R->y[i], R->z[i], R->P->r0, R->P->E0, R->P->xy, R->P->ntri, R->P->tri,
                           R->P->nbr, R->P->er, R->P->ez, &tri, e);
            double f = f0 * R->q[i] / (double)orc_species_mass_ld(R->species[i]);
            ax += f * e[0];
            ay += f * e[1];
            az += f * e[2];
        }
        const double dl = R->dtl[R->level[i]];
        const double s = fmin(fmax(R->tvel + dl - R->tinj[i], 0.0), dl);
        R->vx[i] += ax * s;
        R->vy[i] += ay * s;
        R->vz[i] += az * s;
    }
    R->tvel += R->dtl[R->M];
}

static void ss_crossover(substep_run* R, int L) {
    for (int i = 0; i < R->n; i++) {
        if (R->tinj[i] > R->tpos) continue;
        if (R->level[i] == L || R->level[i] == L + 1)
            R->level[i] = orc_substep_level_of(R->T, R->M, L, R->level[i], R->x[i], R->y[i], R->z[i], R->tpos,
                                               R->tinj[i]);
    }
}

static void ss_K(substep_run* R, int L, int top);
static void ss_D(substep_run* R, int L, int toggle) {
    if (L != R->M) {
        ss_D(R, L + 1, 0);
        ss_K(R, L + 1, L + 1);
        ss_D(R, L + 1, 1);
        if (toggle) ss_crossover(R, L);
    } else {
        ss_drift(R);
    }
}
static void ss_K(substep_run* R, int L, int top) {
    if (L != R->M) {
        ss_D(R, L + 1, 0);
        ss_K(R, L + 1, top);
        ss_D(R, L + 1, 1);
    } else {
        ss_kick(R, top);
    }
}

/* nsteps of LeapfrogDKDSubsteps with the FP64 direct force (the substep
 * method is the same force); level[] in/out (start at M); acc3 receives the
 * last force. */
void orc_leapfrog_substeps_f64(const orc_lf_params* P, const orc_substep_level* T, int M, int n, double* x,
                               double* y, double* z, double* vx, double* vy, double* vz, const double* q,
                               int* status, const int* species, const double* tinj, int* level, double t0,
                               double dt, int nsteps, double* acc3) {
    substep_run R;
    R.P = P;
    R.T = T;
    R.M = M;
    R.n = n;
    R.x = x; R.y = y; R.z = z; R.vx = vx; R.vy = vy; R.vz = vz;
    R.q = q; R.tinj = tinj; R.status = status; R.species = species; R.level = level;
    R.dtl = (double*)malloc(sizeof(double) * (size_t)(M + 1));
    R.dtl[0] = dt;
    for (int l = 1; l <= M; l++) R.dtl[l] = R.dtl[l - 1] / 3.0;
    R.eff = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    R.acc3 = acc3;
    for (int k = 0; k < nsteps; k++) {
        R.tpos = R.tvel = t0 + k * dt;
        ss_D(&R, 0, 0);
        ss_K(&R, 0, 0);
        ss_D(&R, 0, 1);
    }
    free(R.dtl);
    free(R.eff);
}
