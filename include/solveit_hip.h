/*
 * solveit_hip.h -- C ABI of the MI355X force engine (libsolveit_hip.so).
 *
 * Plain C: opaque handles, POD structs, raw pointers and sizes, int status
 * codes (0 = ok, <0 = error; shp_last_error() returns the message of the last
 * failing call on the calling thread).
 *
 * Two families of entry points:
 *   - *_host: take the reference's host AoS arrays (vec4<float> position
 *     (x,y,z,q), velocity (vx,vy,vz,species), acceleration (ax,ay,az,status),
 *     particles.h:99-101) and return when the result is in host memory.  These
 *     are what a C++ ForceCalculationMethod/Integrator subclass calls.
 *   - *_dev: take device pointers to the engine's FP64 SoA state and are
 *     stream-ordered on the handle's stream (shp_sync() waits).
 *
 * Status values follow structures.h:206-211: 1 inbound (target and source),
 * 2 affects-acceleration (source only), 0 out of bounds, -1 crashed.
 * Species codes follow structures.h:224-233 (0 monomer+, 1 dimer+, 5 BF4-, ...).
 */
#ifndef SOLVEIT_HIP_H
#define SOLVEIT_HIP_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SHP_OK 0
#define SHP_ERR_INVALID (-1)   /* bad argument */
#define SHP_ERR_HIP (-2)       /* HIP runtime error */
#define SHP_ERR_NOMEM (-3)     /* device allocation failed */
#define SHP_ERR_UNSUPPORTED (-4)

/* Arithmetic contract of a force evaluation. */
#define SHP_PREC_REF 0   /* reference contract: float displacements, double sums */
#define SHP_PREC_FP64 1  /* FP64 state, FP64 pair arithmetic */
#define SHP_PREC_FP32C 2 /* FP32 pair arithmetic, compensated per-target sums */

/* Symmetry (SymmetryType, structures.h:201-205). */
#define SHP_SYM_NONE 0
#define SHP_SYM_ROTATIONAL 1    /* nsec sectors about z (reference: 4) */
#define SHP_SYM_TRANSLATIONAL 2 /* explicit lattice images */

typedef struct shp_context* shp_handle;

typedef struct shp_symmetry {
    int type;              /* SHP_SYM_* */
    int nsec;              /* rotational: number of sectors (>= 1) */
    int nimg;              /* translational: number of lattice points */
    const int* lattice;    /* host, 3*nimg ints: image = lattice * domain */
    double domain[3];      /* translational cell size (m) */
} shp_symmetry;

/* Device-resident FP64 particle state (SoA, length n). */
typedef struct shp_particles {
    int n;
    const double* x;
    const double* y;
    const double* z;
    const double* q;       /* charge in units of e */
    const int* status;     /* ParticleStatus */
    const int* species;    /* ParticleSpecies -> mass */
} shp_particles;

/* ---- lifetime ---- */
int shp_version(void);
const char* shp_last_error(void);
int shp_create(int device, shp_handle* out);
int shp_destroy(shp_handle h);
int shp_sync(shp_handle h);
/* hipStream_t of the handle (as void*); all *_dev calls are ordered on it. */
void* shp_stream(shp_handle h);

/* ---- constants of the reference (constants.h:28-47, particles.h:67-87) ---- */
double shp_k_const(void);          /* K_CONST */
double shp_softening(void);        /* SOFTENING (eps^2, m^2) */
double shp_species_mass(int species);

/*
 * Direct softened Coulomb sum, reference contract, host AoS in/out.
 * Replaces DirectForceFunction::direct / direct_rotational / direct_translational
 * as driven by DirectForceCalculation::updateAcceleration
 * (direct_force_method.cpp:153-190, :347-428; direct_force_method_rotational.cpp:28-94;
 * direct_force_method_translational.cpp:28-367).
 * Targets [i0,i1) with status 1; sources [j0,j1) with status 1 or 2.
 * out3[3*(i-i0)+c] receives the double acceleration before the reference's
 * rounding into float (0 for non-targets).  precision: SHP_PREC_REF or
 * SHP_PREC_FP32C (the latter on the same float inputs).
 */
int shp_direct_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4,
                    int n, int i0, int i1, int j0, int j1, int precision,
                    const shp_symmetry* sym, double* out3);

/*
 * Direct sum on the device-resident FP64 state: every status-1 particle is a
 * target, every status-1/2 particle a source.  ax/ay/az (device, length n)
 * receive K*q_i/m_i * sum_j q_j (r_i - r_j)/(|r_i - r_j|^2 + eps)^1.5, zero for
 * non-targets.  precision: SHP_PREC_FP64 or SHP_PREC_FP32C.
 */
int shp_direct_dev(shp_handle h, const shp_particles* p, int precision,
                   const shp_symmetry* sym, double* ax, double* ay, double* az);

/* The same as shp_direct_dev, allocating nothing: every device buffer it
 * needs (staged sources, target order, split partials, sort and selection
 * storage) is carved from a per-device scratch array of
 * SHP_NOALLOC_SCRATCH_BYTES that the library holds from shp_create until the
 * process ends.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  n is at most SHP_NOALLOC_MAX_N and
 * a symmetry has at most 1024 images; larger inputs return SHP_ERR_INVALID.
 * Calls that overlap on one GPU share its scratch and race: serialise them.
 * The results equal shp_direct_dev's bit for bit. */
#define SHP_NOALLOC_MAX_N 262144
#define SHP_NOALLOC_SCRATCH_BYTES (64u << 20)
int shp_direct_dev_noalloc(shp_handle h, const shp_particles* p, int precision,
                           const shp_symmetry* sym, double* ax, double* ay, double* az);

/* Kernel time of the last shp_direct_dev/shp_direct_host force kernel (ms),
 * measured with HIP events on the handle's stream (for shp_fmm_dev: the span
 * from P2M to the last leaf kernel). */
double shp_last_kernel_ms(shp_handle h);

/*
 * Fast multipole evaluation of the same softened sum on the device-resident
 * FP64 state (FP32 expansions and pairs, compensated per-target sums).
 * Replaces MultipoleMethodForceCalculation::updateAcceleration and the passes
 * it drives (multipole_method_force_calculation.cpp:148-522: P2P, P2M, M2M,
 * M2L, L2L, L2P of fmm_cpukernel.cpp:103-1317).  order: expansion order p
 * (1..12); leaf_level: uniform leaf level of the octree (0 = choose it: let
 * Lr be the smallest level whose mean leaf occupancy is <= leaf_size; for
 * orders 4, 6, 8 and 10 with ring 2 the levels max(2, Lr-2) .. Lr+1 are
 * compared by an estimate of the pass times from exact counts of near-field
 * particle pairs and far-band box pairs (a function of the ion positions
 * only), otherwise Lr is used); ring: near-field width in leaf cells (1 or 2;
 * the far field starts beyond it).  The reference chooses its depth from N
 * alone (setOptimumLevel, multipole_method_force_calculation.cpp:835-881).
 */
int shp_fmm_dev(shp_handle h, const shp_particles* p, int order, int leaf_level, int leaf_size,
                int ring, double* ax, double* ay, double* az);

/* FMM with symmetry images, the FMM counterpart of shp_direct_dev's
 * rotational and translational sums: replaces the rotational and
 * translational multipole methods (multipole_method_force_calculation_rotational.cpp:9-333,
 * multipole_method_force_calculation_translational.cpp:9-520).  The images
 * (rotations by 2 pi k / nsec about z, k = 1..nsec-1, or the lattice points
 * other than (0,0,0), which must be present) enter the tree as source-only
 * ions (status 2), so every image of every ion is a source and the ions
 * themselves are the targets (self excluded, images of self included, as in
 * the reference's image sums).  sym NULL or SHP_SYM_NONE = shp_fmm_dev. */
int shp_fmm_sym_dev(shp_handle h, const shp_particles* p, const shp_symmetry* sym, int order,
                    int leaf_level, int leaf_size, int ring, double* ax, double* ay, double* az);

/* The FMM on the reference's host AoS arrays, the drop-in for
 * MultipoleMethodForceCalculation::updateAcceleration
 * (multipole_method_force_calculation.cpp:148-182) and its rotational and
 * translational siblings behind ForceCalculationMethod
 * (force_calculation_methods.h:28-87): the ions [i0, i1) are the sources
 * (status 1 or 2) and targets (status 1), as [firstParticle, lastParticle);
 * order, leaf_level, leaf_size, ring and sym as for shp_fmm_sym_dev (sym NULL
 * = no symmetry).  out3[3*(i-i0)+c] receives the acceleration (m/s^2) as
 * double, zero rows for non-targets; the host arrays are not modified (the
 * reference permutes and restores them, multipole_method_force_calculation.cpp:1086-1242).
 * Waits for the result. */
int shp_fmm_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n,
                 int i0, int i1, int order, int leaf_level, int leaf_size, int ring,
                 const shp_symmetry* sym, double* out3);

/* The reference's own FMM operators, the drop-in for
 * MultipoleMethodForceCalculation::updateAcceleration with its spherical-
 * harmonic expansions (multipole_method_force_calculation.cpp:148-522;
 * p2m_kernel, m2m_kernel, m2l_kernel, l2l_kernel, l2p_kernel and rotation,
 * fmm_cpukernel.cpp:10-1317; the tables of MultipoleMethodConstants::precalc,
 * multipole_method_constants.cpp:40-226): expansions of 1/r translated by
 * rotation, shift along z and rotation back, the softened centre distance in
 * M2L (fmm_cpukernel.cpp:832), P2P over the 27 neighbour leaves.  order is
 * numExpansions (1..10, setNumExpansions); max_level is setMaxLevel's value
 * (-2: setOptimumLevel's choice from the largest domain's count).  The tree
 * is the reference's: the domain grid of computeDomainBounds, maxLevel levels
 * in every domain, float leaf placement (morton,
 * multipole_method_force_calculation.cpp:976-1057); above the domain level,
 * where the reference groups several domains into ternary supercells, octree
 * levels of one enclosing cube.  Computed in FP64 (the translations as dense
 * operators on the FP64 matrix cores).  Ranges and out3 as shp_fmm_host.
 * sym NULL or SHP_SYM_NONE: updateAccelerationNoSym; SHP_SYM_ROTATIONAL with
 * nsec 4: updateAccelerationRotational (multipole_method_force_calculation_
 * rotational.cpp:9-331, fmm_cpukernel_rotational.cpp): the ions in the first
 * quadrant of a domain column centred on the z axis (computeDomainBounds-
 * Rotations, domain_bounds_functions.cpp:317-500), the other quadrants their
 * images, mapped back by getInteractionListRotations (:568-1438).
 * SHP_SYM_TRANSLATIONAL: updateAccelerationTranslational (multipole_method_
 * force_calculation_translational.cpp:9-331, fmm_cpukernel_translational.cpp):
 * one cubic periodic cell of side domain[0] (= domain[1] = domain[2]) centred
 * on the origin, a ternary tree of levels 1 .. maxLevel (1..6), lists wrapped
 * periodically (getInteractionListTranslations :1179-1834), so every target
 * sees the sources of the 27 cells at shifts {-1, 0, 1}^3; the lattice must be
 * those 27 points (any order).  Its P2P adds direct_translational over each
 * wrapped neighbour leaf pair (the reference's p2p_kernel_translations passes
 * the thread's box range as the ion range and adds nothing, :305-336).
 * SHP_ERR_UNSUPPORTED when a source lies outside the domain grid or the cell
 * (the reference's outlier passes), maxLevel < 2 (translational: < 1; its
 * single-box and supercell paths), other sector counts or image sets, and for
 * a rotational grid with more than one domain in x or y (the reference's
 * quadrant map works in domain-local x, y).  Waits. */
int shp_fmm_sph_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n,
                     int i0, int i1, int order, int max_level, const shp_symmetry* sym, double* out3);
/* Summary of the last shp_fmm_sph_host call, 12 values: [0] maxLevel, [1]
 * levels above the domain level, [2..4] domains per axis, [5] leaf boxes,
 * [6] M2L box pairs, [7] P2P box pairs, [8] sources, [9] boxes of all levels
 * from level 2, [10] P2P ion pairs (each leaf's ions times its neighbour
 * list's ions), [11] reserved.  No counterpart in the reference (its
 * timing file records the level and domain counts). */
int shp_fmm_sph_stats(shp_handle h, double* out);

/* M2L expansion orders the engine uses on the Chebyshev shells 3, 4 and 5 of
 * the far band for an FMM of this order and near-field width (sym_type
 * SHP_SYM_ROTATIONAL: the rotational variant of shp_fmm_sym_dev, which keeps
 * the full order on every shell); out3 = {0, 0, 0} when the order has no
 * compiled term list and the table-driven M2L at the single order is used.
 * A pure function (no device work).  No counterpart in the reference. */
int shp_fmm_shell_orders(int order, int ring, int sym_type, int* out3);

/* FmmType (structures.h:283-288) of later FMM evaluations on this handle
 * (shp_fmm_dev, shp_fmm_sym_dev, shp_fmm_host, the FMM leapfrog):
 *  - SHP_FMM_FMM (the default): M2L, L2L and L2P, as above;
 *  - SHP_FMM_TREECODE: no local expansions; every target evaluates the
 *    multipoles of its far boxes at every level directly (M2P,
 *    fmm_cpukernel.cpp:1319-1533), with the same near field;
 *  - SHP_FMM_FMM_OUTLIERS / SHP_FMM_TREECODE_OUTLIERS: the ions that
 *    DomainBoundsFunction::isOutlier flags (deviation from the median of at
 *    least 3 x 1.4826 median absolute deviations on some axis,
 *    domain_bounds_functions.cpp:16-93) are held out of the tree; they get the
 *    tree's field by M2P from their far boxes plus the ions of their near
 *    cells (m2outlier_p, p2outlier_p, fmm_cpukernel.cpp:2282-2776), and every
 *    target gets the outliers' field by a direct FP64 sum (outlier_p2p,
 *    :2230-2280).  Without symmetry only (SHP_ERR_UNSUPPORTED otherwise).
 * M2P evaluates expansions of order <= 10; a larger order in these modes
 * returns SHP_ERR_UNSUPPORTED. */
enum { SHP_FMM_TREECODE = 0, SHP_FMM_TREECODE_OUTLIERS = 1, SHP_FMM_FMM = 2, SHP_FMM_FMM_OUTLIERS = 3 };
int shp_fmm_set_type(shp_handle h, int fmm_type);
/* The outlier flags (1 = outlier) of the outlier modes for device arrays
 * (flags: n device ints), and their number (*nout, may be null). */
int shp_fmm_outliers_dev(shp_handle h, const shp_particles* pt, int* flags, int* nout);

/* Leaf level used by the last shp_fmm_dev call on this handle. */
int shp_fmm_last_leaf_level(shp_handle h);

/* Measurements of the last shp_fmm_dev call (waits for it), 11 values:
 * out[0..4] = ms of P2M+M2M, translation tables, M2L, L2L, L2P+P2P;
 * out[5] = M2L box pairs; out[6] = P2P particle pairs; out[7] = leaf level;
 * out[8] = M2L multiply-adds (expansion terms of the box pairs that exist);
 * out[9] = ms of the first M2L kernel (output degrees 0..4, the longest);
 * out[10] = its multiply-adds.  No counterpart in the reference (its
 * updateAcceleration has no pass timers). */
int shp_fmm_stats(shp_handle h, double* out);

/* Sums of the shp_fmm_stats values over every shp_fmm_dev call (including the
 * calls inside shp_leapfrog_run) since the last reset: out[0..10] as above,
 * summed; out[11] = number of calls.  out may be NULL; reset != 0 clears the
 * sums after reading.  Waits for the last call.  Measurement only: no
 * counterpart in the reference. */
int shp_fmm_timing(shp_handle h, double* out, int reset);

/* The leaf-level choice of the last shp_fmm_dev call with leaf_level 0:
 * one row per candidate level, (level, near-field particle pairs, far-band
 * box pairs of that level, estimated ms of the level-dependent passes), in
 * increasing level order; at most cap rows are written to out (4 doubles
 * each).  Returns the number of candidate rows (0 when the level was given or
 * the rule was used), or a negative error code.  No counterpart in the
 * reference. */
int shp_fmm_level_estimates(shp_handle h, double* out, int cap);

/*
 * Barnes-Hut evaluation, reference (FP32) arithmetic.  Replaces
 * BarnesHutForceCalculation::updateAcceleration (barnes_hut_force_calculation.cpp:574-715:
 * computeWorldLimits :456-522, generateTree :258-331, computeCellProperties
 * :337-390, computeForceFromTree :417-450, barnesHutKernelCalculation :539-569),
 * DomainBoundsFunction::computeDomainBounds (domain_bounds_functions.cpp:121-316)
 * and its rotational (barnes_hut_force_calculation_rotational.cpp:28-135, four
 * quadrants: sym->nsec must be 4) and translational
 * (barnes_hut_force_calculation_translational.cpp:190-280, lattice images in
 * the given order) variants.  The reference's grid of root cells chained as
 * siblings, one ion per leaf (ions that agree to 42 levels below a root cell
 * share a bucket leaf), monopole at the centre of charge, accepted iff
 * theta (d - delta) > repLength.  Sources: status 1 or 2; targets: status 1.
 */
int shp_bh_dev(shp_handle h, const shp_particles* p, float theta, const shp_symmetry* sym,
               double* ax, double* ay, double* az);
/* The same on the reference's host AoS arrays for ions [i0, i1) (sources and
 * targets); out3[3*(i-i0)+c] receives the float acceleration the reference
 * accumulates, as double (0 for non-targets). */
int shp_bh_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4, int n,
                int i0, int i1, float theta, const shp_symmetry* sym, double* out3);

/* Root-layout options of the handle's Barnes-Hut evaluations, the setters of
 * BarnesHutForceCalculation (force_calculation_methods.h:222-329):
 * setDomainToggle (0 cubic cells, the default; 1 rectangular prism),
 * setBoxDimVals (force_box_dims = 1 and the cell counts), setSubcellBlockSize
 * (2 only) and setOutlierToggle (0 only; other values return
 * SHP_ERR_UNSUPPORTED).  walk (no counterpart in the reference) picks the
 * GPU walk: 0 by the number of targets (the default), 1 each lane on its
 * own, 2 the lanes of a wave together; all give the same bits. */
typedef struct shp_bh_options {
    int domain_toggle;
    int force_box_dims;
    int box_dims[3];
    int subcell_block;
    int outlier_toggle;
    int walk;
} shp_bh_options;
int shp_bh_set_options(shp_handle h, const shp_bh_options* o);

/* Summary of the last Barnes-Hut evaluation (waits for it): out[0] root cells
 * of the grid, [1] tree nodes (non-empty cells), [2] deepest level below a
 * root cell, [3] bucket leaves, [4] accepted cells, [5] ion terms (both
 * summed over targets and images), [6] sources placed in the last root cell
 * because no cell matched, [7..9] cells per axis, [10..12] location of the
 * first root cell, [13..15] root cell dimensions, [16] opened cells (summed
 * over targets and images).  Has no counterpart in the reference (its timing
 * record is BarnesHutForceCalculation's timeInfoFile). */
#define SHP_BH_NSTATS 17
int shp_bh_stats(shp_handle h, double* out);

/* ---- energies (FP64) ---- */

/* Kinetic and potential energy (J) of the ions with status 1 or 2 (the caller
 * marks the injected set): out[0] = sum 0.5 amu m |v|^2; out[1] = 1/2 sum over
 * ordered pairs i != j of e^2/(4 pi eps0) q_i q_j / sqrt(r^2 + eps^2), an FP64
 * direct pair sum.  Replaces ParticleSystem::computeChargedKineticEnergy and
 * computeChargedElectricPotentialEnergy (particles.cpp:535-578) as summed and
 * halved by Integrator::updateEnergyIntervalData (integration_schemes.cpp:799-835).
 * Waits for the result (out is host memory). */
int shp_energy_dev(shp_handle h, const shp_particles* p, const double* vx, const double* vy,
                   const double* vz, double* out);

/* Field energy: fe[i] += -e q_i (sum_k E_z dz/ns + sum_k (E_x dx/ns + E_y dy/ns))
 * along the path (x0, y0, z0) -> (x0, y0, z) -> (x, y, z) with left Riemann sums
 * of ns = nsteps_arr[i] (or nsteps when nsteps_arr is NULL) points per leg, E
 * from the mesh of shp_efield_setup.  Replaces
 * ParticleSystem::updateChargedFieldEnergy (particles.cpp:598-638).  Device
 * arrays; stream-ordered. */
int shp_field_energy_dev(shp_handle h, int n, const double* x0, const double* y0,
                         const double* z0, const double* x, const double* y, const double* z,
                         const double* q, const int* nsteps_arr, int nsteps, double* fe);

/* ---- accuracy metric ---- */

/* Accuracy of an acceleration field against a reference one, as
 * Integrator::compareAcceleration computes it (integration_schemes.cpp:374-470),
 * in FP64: over the ions with status 1, rel_i = |a_i - r_i|^2 / |r_i|^2 (0 when
 * both are 0); out[0] = number of such ions, out[1] = sqrt(mean rel) (NaN
 * when there are none, as in the reference), out[2] = sqrt(max rel), out[3] =
 * the first index of the maximum (0 when every rel is 0), i.e. the values of
 * ForceCalculationMethod::writeL2NormInfo.  Device arrays of length n; a
 * deterministic two-level reduction; waits for the result (out is host memory). */
int shp_compare_accel(shp_handle h, int n, const int* status, const double* ax, const double* ay,
                      const double* az, const double* rx, const double* ry, const double* rz,
                      double* out);

/* ---- background E-field (pwl on a scattered (r, z) triangulation) ---- */

/* Host arrays in the reference's layout (InterpolationScheme members,
 * interpolation_scheme.h:44-66): nodes xy[2*nnodes] (r/r0, z/r0), er/ez per node
 * (field in units of E0), triangles tri[3*ntri] (0-based), neighbours
 * nbr[3*ntri] (neighbour k across the edge opposite vertex k, <0 on the hull). */
typedef struct shp_field_mesh {
    int nnodes;
    const double* xy;
    const double* er;
    const double* ez;
    int ntri;
    const int* tri;
    const int* nbr;
    double r0;   /* InterpolationScheme::r0 (m) */
    double E0;   /* InterpolationScheme::E0 (V/m) */
} shp_field_mesh;

/* Uploads the mesh and builds the point locator.  Replaces the one-time
 * InterpolationScheme::initializeSingleEmitterInterpolationScheme
 * (interpolation_scheme.cpp:152-191). */
int shp_efield_setup(shp_handle h, const shp_field_mesh* mesh);

/* Reads the reference's field files (InterpolationScheme::fillInterpolationScheme
 * and fillTriangleScheme, interpolation_scheme.cpp:66-172; line counts as
 * returnSize, auxiliary_functions.cpp:39-53): Fields_*.csv with x, y (units
 * of r0), Er, Ez (units of E0) per line, Triangles_*.csv with three 0-based
 * node ids and three neighbours per line, fields split on ',' and parsed with
 * atof / atoi.  Fills *mesh (r0, E0 as given) with host arrays owned by the
 * library; shp_field_mesh_free releases them.  A file that does not open is
 * an error (the reference prints and exits).  No device work. */
int shp_field_mesh_load_csv(const char* fields_path, const char* triangles_path, double r0, double E0,
                            shp_field_mesh* mesh);
void shp_field_mesh_free(shp_field_mesh* mesh);

/* Field at n device points (FP64): E = E0 (Er cos(phi), Er sin(phi), Ez) with
 * (Er, Ez) interpolated at (rho/r0, z/r0), zero outside z/r0 in [0, 140].
 * Replaces InterpolationScheme::computeEfieldSingleEmitter
 * (interpolation_scheme.cpp:200-271) and the walk of
 * pwl_interp_2d_scattered_value (pwl_interp_2d_scattered.c:887-956, :1762-1990).
 * tri_out (device, may be NULL) receives the triangle used. */
int shp_efield_dev(shp_handle h, int n, const double* x, const double* y, const double* z,
                   double* ex, double* ey, double* ez, int* tri_out);

/* ---- leapfrog DKD on the device-resident FP64 state ---- */

typedef struct shp_state {
    int n;
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    const double* q;
    int* status;            /* updated by boundary conditions */
    const int* species;
    const double* t_inj;    /* injection time (s), ascending */
} shp_state;

#define SHP_METHOD_DIRECT 0
#define SHP_METHOD_FMM 1
#define SHP_METHOD_BH 2

typedef struct shp_step_opts {
    int method;          /* SHP_METHOD_* */
    int precision;       /* SHP_PREC_FP64 or SHP_PREC_FP32C */
    shp_symmetry sym;    /* force symmetry; also drives folding / wrapping */
    int use_efield;      /* add the background field (needs shp_efield_setup) */
    double bc_space[4];  /* |x|, |y|, |z|, rho limits (m); 0 disables each */
    double bc_time;      /* max time since injection (s); 0 disables */
    int fmm_order;       /* FMM expansion order p */
    int leaf_size;       /* FMM / Barnes-Hut leaf capacity */
    double theta;        /* Barnes-Hut opening angle */
    int fmm_ring;        /* FMM near-field width in leaf cells (1 or 2) */
    int leaf_level;      /* FMM uniform leaf level (0 = from leaf_size) */
} shp_step_opts;

/*
 * Runs nsteps of LeapfrogDKD::runSimulation (integration_schemes.cpp:856-891):
 * per step drift(dt/2) -- force -- E-field -- kick(dt) -- drift(dt/2), with the
 * injection clamps of propagatePositionNew/propagateVelocityNew
 * (particles.cpp:693-786), the symmetry folding of enforceSymmetryConditions
 * (particles.cpp:199-340) and the boundary conditions of
 * enforceBoundaryConditions (particles.cpp:73-177).  Ions with t_inj later than
 * the drift's end time are outside the active range (findMostRecentInjected,
 * particles.cpp:1328-1359).  ax/ay/az (device, length n) hold the last
 * Coulomb acceleration.  Stream-ordered.
 */
int shp_leapfrog_run(shp_handle h, shp_state* s, const shp_step_opts* o, double t0, double dt,
                     int nsteps, double* ax, double* ay, double* az);

/* LeapfrogDKDSubsteps (integration_schemes.cpp:899-1529, class at
 * integration_schemes.h:104-161): multi-rate DKD with levels 0..max_level of
 * time step dt_l = dt_{l-1} / 3.  levels[l] is the level's row of the
 * reference's substepBoundaryConditions table after setSubstepTimes
 * (:1375-1460): the boundary-condition type (BoundaryConditionOptions: 0 none,
 * 1 space, 2 time, 3 space and time), the spatial limits (|x|, |y|, |z|, rho;
 * 0 = none) and the time-since-injection limit.  level (device, n ints, in
 * and out) is particleLevels: start every ion at max_level
 * (initializeParticleLevels :1481-1489).  Per step all injected ions drift
 * together in steps of dt_max / 2; each of the 3^max_level finest kicks sets
 * the status of the injected ions from their levels (setParticleStatus
 * :1193-1211), evaluates the force (o when the kick's top level is 0, sub
 * otherwise, NULL = o) and kicks the status-1 ions by their own level's dt
 * (variableTimestepKick :1216-1250); crossoverParticles (:1117-1127)
 * re-levels ions by the table after each level's second drift.  Options,
 * field, folding and boundary conditions as for shp_leapfrog_run; max_level
 * at most 7.  Stream-ordered. */
typedef struct shp_substep_level {
    int bc_type;
    double bc_space[4];
    double bc_time;
} shp_substep_level;
int shp_leapfrog_substeps_run(shp_handle h, shp_state* s, const shp_step_opts* o, const shp_step_opts* sub,
                              const shp_substep_level* levels, int max_level, int* level, double t0,
                              double dt, int nsteps, double* ax, double* ay, double* az);

/* Resident state on the reference's host arrays, for an Integrator subclass
 * (HipLeapfrogDKD) that keeps the plume on the GPU between steps:
 *  - shp_state_upload_host copies position (x, y, z, q), velocity (vx, vy,
 *    vz, species), acceleration (ax, ay, az, status) and timeInjected
 *    (particles.h:99-104) for ions [0, n) into the handle's FP64 state;
 *  - shp_state_run advances it by nsteps of shp_leapfrog_run from t0;
 *  - shp_state_download_host writes positions, velocities, the last Coulomb
 *    acceleration and the status back as float (q and species unchanged).
 * The upload and download wait for their copies. */
int shp_state_upload_host(shp_handle h, const float* pos4, const float* vel4, const float* acc4,
                          const float* t_inj, int n);
int shp_state_run(shp_handle h, const shp_step_opts* o, double t0, double dt, int nsteps);
int shp_state_download_host(shp_handle h, float* pos4, float* vel4, float* acc4);
/* shp_leapfrog_substeps_run on the resident state, with the ion levels kept
 * in the handle between calls (reset_levels != 0 sets every ion to
 * max_level first, as initializeParticleLevels does). */
int shp_state_run_substeps(shp_handle h, const shp_step_opts* o, const shp_step_opts* sub,
                           const shp_substep_level* levels, int max_level, int reset_levels, double t0,
                           double dt, int nsteps);

/* ---- writers of the resident state (the reference's binary formats) ----
 * While shp_state_run / shp_state_run_substeps drive the leapfrog, the engine
 * keeps each ion's background field at its last kick (the reference's
 * Efield2, particles.cpp:508-512) and queues the boundary-condition events
 * that the reference writes from enforceBoundaryConditions
 * (particles.cpp:73-166); an upload resets both. */

/* SaveDataOptions (structures.h:244-247). */
enum { SHP_SAVE_ORIGINAL = 0, SHP_SAVE_NEW_ORIGINAL = 1 };

/* Integrator::saveData (integration_schemes.cpp:263-331): the ions of
 * [first, last) with status 1 or 2, in index order, in the record layout of
 * `format` (original: x y z vx vy vz, Efield1, Efield2 as floats, species,
 * index, lastParticle = last as 32-bit integers; newOriginal: a header
 * 0xffffffff, lastParticle, current_time, then index, species and the twelve
 * floats per ion).  Efield1 is the Coulomb field AMU/e * a * m / q of the last
 * force evaluation, rounded as the reference rounds it
 * (integration_schemes.cpp:350-355).  The snapshot is packed on the GPU, copied
 * to pinned host memory on a second stream and written to `out` by a host
 * thread: the call returns once the pack is enqueued, and `out` must not be
 * touched or closed before shp_state_flush returns.  Two snapshots can be in
 * flight; a third call waits for the oldest. */
int shp_state_save(shp_handle h, int format, int first, int last, float current_time, FILE* out);
/* Integrator::saveFullCurrentConditions (integration_schemes.cpp:197-258):
 * every resident ion, whatever its status, in the same layouts, with
 * lastParticle = last_particle.  Asynchronous like shp_state_save. */
int shp_state_save_full(shp_handle h, int format, int last_particle, float current_time, FILE* out);
/* Waits until every snapshot is written; returns the first write failure. */
int shp_state_flush(shp_handle h);
/* ChargedParticles::writeToEventsFile (particles.cpp:666-684) for the events
 * queued since the last call: time (the drift's end), x y z, vx vy vz as
 * floats, index, species, event (0, outOfBounds) as 32-bit integers, in the
 * order the reference writes them (by drift, spatial limits before the time
 * limit, then index).  Synchronous; *nevents (may be null) gets the count. */
int shp_state_events(shp_handle h, FILE* out, int* nevents);
/* Efield1 and Efield2 of every resident ion into the reference's
 * vec3<float> arrays (ChargedParticles::Efield1 / Efield2, particles.h:108-109). */
int shp_state_fields_host(shp_handle h, float* efield1, float* efield2);

/* Kernel time (ms) of the last leapfrog update kernel. */
double shp_last_update_ms(shp_handle h);

/* ---- derivatives (FP64; no counterpart in the reference) ---- */

/* Force JVP: tangent of the FP64 direct acceleration (shp_direct_dev,
 * SHP_PREC_FP64) for position tangents (dx, dy, dz) and charge tangents dq
 * (any may be NULL = 0):
 *   da_i = k_i sum_j q_j H(r_ij)(dx_i - dx_j) + k_i sum_j dq_j G(r_ij) + K dq_i/m_i sum_j q_j G(r_ij),
 * G(r) = r s^-3/2, H(r) = I s^-3/2 - 3 r r^T s^-5/2, s = |r|^2 + eps^2: an
 * N-body sum with a rank-2 kernel.  Zero for non-targets. */
int shp_force_jvp_dev(shp_handle h, const shp_particles* p, const double* dx, const double* dy,
                      const double* dz, const double* dq, double* oax, double* oay, double* oaz);

/* Force VJP: for an acceleration cotangent (bax, bay, baz) writes (or, with
 * accumulate != 0, adds) the cotangents of positions (xb, yb, zb) and, when qb
 * is not NULL, of charges. */
int shp_force_vjp_dev(shp_handle h, const shp_particles* p, const double* bax, const double* bay,
                      const double* baz, double* xb, double* yb, double* zb, double* qb,
                      int accumulate);

/* VJP of one DKD leapfrog step (shp_leapfrog_run with the FP64 direct force,
 * all ions injected, no background field or boundary conditions; status 1
 * ions drift and kick, status 2 drift, others are frozen) taken from the
 * state s at the start of the step: for cotangents of the end state
 * (xb1 .. vzb1) writes the cotangents of the start state (xb0 .. vzb0) and
 * adds the charge cotangent into qb.  Checkpointed reverse sweeps call it once
 * per step from the last step back. */
int shp_leapfrog_step_vjp_dev(shp_handle h, const shp_state* s, double dt, const double* xb1,
                              const double* yb1, const double* zb1, const double* vxb1,
                              const double* vyb1, const double* vzb1, double* xb0, double* yb0,
                              double* zb0, double* vxb0, double* vyb0, double* vzb0, double* qb);

/* ---- batched plumes (no counterpart in the reference, which runs one plume per process) ---- */

/* A batch is the concatenated device SoA state of ninst independent plumes:
 * instance k holds ions [offsets[k], offsets[k+1]) and its ions interact only
 * with each other.  The plan holds the work list of a persistent direct-sum
 * launch (tiles of targets, largest instances first, pulled from an atomic
 * counter) for precision SHP_PREC_FP32C or SHP_PREC_FP64. */
typedef struct shp_batch_plan_s* shp_batch_plan;
int shp_batch_create(shp_handle h, int ninst, const int* offsets, int precision,
                     shp_batch_plan* out);
int shp_batch_destroy(shp_batch_plan b);
/* Times the candidate tile shapes (workgroup size x targets per lane) on the
 * given state and keeps the fastest; ms_out (may be NULL) receives the
 * average time of each candidate, chosen its index. */
int shp_batch_tune(shp_batch_plan b, const shp_particles* all, double* ms_out, int* chosen);
int shp_batch_set_config(shp_batch_plan b, int config);
/* Accelerations of every instance (one persistent launch); stream-ordered. */
int shp_batch_forces(shp_batch_plan b, const shp_particles* all, double* ax, double* ay,
                     double* az);
/* nsteps DKD steps of every instance (shp_leapfrog_run semantics, no
 * symmetry); with use_graph the step is captured once in a HIP graph (time
 * kept in device memory) and replayed.  Stream-ordered. */
int shp_batch_leapfrog(shp_batch_plan b, shp_state* all, const shp_step_opts* o, double t0,
                       double dt, int nsteps, int use_graph, double* ax, double* ay, double* az);
/* VJP of one DKD step of every instance (FP64 force VJP, semantics of
 * shp_leapfrog_step_vjp_dev) from the batch state s at the start of the step.
 * Cotangents are component-major device arrays of 3n doubles (x components,
 * then y, then z); qb (n) accumulates the charge cotangent. */
int shp_batch_step_vjp(shp_batch_plan b, const shp_state* s, double dt, const double* xb1,
                       const double* vb1, double* xb0, double* vb0, double* qb);

#ifdef __cplusplus
}
#endif
#endif /* SOLVEIT_HIP_H */
