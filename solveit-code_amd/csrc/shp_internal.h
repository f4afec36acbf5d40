// Internal definitions shared by the engine's HIP translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/solveit_hip.h"

namespace shp {

// Physical constants of the reference (constants.h:28-47).  K is evaluated on
// the host exactly as the reference does (float e*e over a long double
// denominator) and handed to kernels as a double argument.
double k_const_host();
constexpr double kSoftening = 6.4e-19;  // eps^2 in m^2

// Mass in amu per ParticleSpecies code (particles.h:67-87, with droplet -> 0
// as ChargedParticles::getMass returns, particles.cpp:1081-1089).
__host__ __device__ inline double species_mass(int s) {
    switch (s) {
    case 0: return 111.0;
    case 1: return 309.0;
    case 2: return 507.0;
    case 4: return 198.0;
    case 5: return 87.0;
    case 6: return 285.0;
    case 7: return 483.0;
    default: return 0.0;
    }
}

// Values reported per FMM call (shp_fmm_stats / shp_fmm_timing).
constexpr int kFmmStats = 11;

void set_error(const char* fmt, ...);

// DomainBoundsFunction::computeDomainBounds (domain_bounds_functions.cpp:121-316)
// without outliers, from the sources' float bounds (mn, mx), the first ion of
// the range (seed) and the number of sources: the first root cell's corner,
// the cell dimensions and the cells per axis (bh.hip).  toggle 0 cubic cells,
// 1 rectangular prism; rot != 0: computeDomainBoundsRotations (:317-500), the
// x, y extent a square about the z axis; forced: NULL or setBoxDimVals' counts.
void domain_bounds_host(const float mn_src[3], const float mx_src[3], const float seed[3], int nsrc,
                        int toggle, int rot, const int* forced, float boxMin[3], float dims[3], int bdv[3],
                        float minO[3], float maxO[3]);

// Leapfrog building blocks shared with the batched driver (leapfrog.hip).
int lf_begin(shp_handle h, const shp_state* s, const shp_step_opts* o, double t0, double dt, int* eff);
int lf_kick_drift(shp_handle h, const shp_state* s, const shp_step_opts* o, double t, const double* tdev,
                  double dt, int next, const double* ax, const double* ay, const double* az, int* eff);
void free_mesh(shp_handle h);
int direct_scratch_setup(int device);
void fmm_release(shp_handle h);
void sph_release(shp_handle h);  // tables of the spherical-harmonic FMM (fmm_sph.hip)

// Writers of the resident state (writers.hip).  One boundary-condition event
// is kEventWords 32-bit words in the layout of writeToEventsFile
// (particles.cpp:666-684): time, x, y, z, vx, vy, vz (float), index, species,
// event (uint32).
constexpr int kEventWords = 10;
struct StateWriter;              // snapshot staging and its host writer thread
void writer_release(shp_handle h);  // waits for pending writes, frees the staging

// Grow-only device workspace slots owned by a context.
enum WsSlot {
    WS_SRC = 0,     // staged sources
    WS_TGT,         // target index list
    WS_KEYS,        // sort keys (in)
    WS_KEYS2,       // sort keys (out)
    WS_VALS2,       // sort values (out)
    WS_PARTIAL,     // split-j partial sums
    WS_COUNTS,      // small device counters
    WS_CUB,         // hipcub temporary storage
    WS_IMAGES,      // per-image transforms
    WS_BBOX,        // bounding-box partials
    WS_FLAGS,       // selection flags
    WS_HOST_IN,     // device copy of host inputs
    WS_HOST_OUT,    // device staging of host outputs
    WS_HOST_IN2,
    WS_HOST_IN3,
    WS_WIDE_F64,    // float AoS widened to FP64 SoA
    WS_WIDE_I32,
    WS_BOX,         // bounding box (6 doubles)
    WS_EFF_STATUS,  // status masked by the active (injected) range
    WS_FMM_KEYS,
    WS_FMM_KEYS2,
    WS_FMM_BKEY,    // box keys of all levels
    WS_FMM_BINT,    // box start/end/first-child of all levels
    WS_FMM_EXP,     // multipole and local coefficients
    WS_FMM_T,       // translation tables
    WS_FMM_GRID,    // dense box-index grids of all levels
    WS_FMM_PART,    // M2L partial sums of the offset chunks
    WS_FMM_OKEY,    // (level, octant) group keys of the M2L boxes (in, out)
    WS_FMM_OVAL,    // box indices grouped by (level, octant) (in, out)
    WS_FMM_SEG,     // group start/end
    WS_FMM_TILES,   // M2L tiles
    WS_BH_SRC,      // Barnes-Hut source list
    WS_BH_SP,       // Barnes-Hut ions in key order
    WS_BH_SI,       // their original indices
    WS_BH_NODES,    // Barnes-Hut tree nodes
    WS_BH_WALK,     // Barnes-Hut walk records (one 32-byte record per node)
    WS_ADJ_U,       // force VJP: scaled cotangents
    WS_ADJ_TMP,     // leapfrog VJP: half-step positions and cotangents
    WS_BATCH_TUNE,  // autotuner scratch outputs
    WS_FMM_SYM_IMG, // FMM symmetry image transforms
    WS_FMM_SYM_D,   // FMM ions with their images: positions, charges, accelerations
    WS_FMM_SYM_I,   // their status and species
    WS_FMM_LCOST,   // FMM leaf-level work estimates
    WS_FMM_UCNT,    // FMM leaf work-unit counts and offsets
    WS_FMM_UNITS,   // FMM leaf work units (leaf, target run)
    WS_FMM_USORT,   // FMM leaf work units: sort keys, values and the longest-first order
    WS_FMM_SPF,     // FMM ions relative to their own leaf centre (FP32)
    WS_FMM_TPRE,    // FMM target flags of the sorted ions and their prefix sums
    WS_FMM_PTS,     // FMM point targets of the M2P pass (treecode: every target)
    WS_FMM_OUTL,    // FMM outlier detection: sort keys, flags, the outlier list
    WS_CMP,         // accuracy metric: per-block partials
    WS_SPH_KEYS,    // spherical FMM: leaf keys and sources (in, out), domain histogram
    WS_SPH_LVL,     // spherical FMM: box keys of all levels, run counts and offsets
    WS_SPH_PART,    // spherical FMM: sources in key order and their leaf boxes
    WS_SPH_EXP,     // spherical FMM: multipole and local expansions of all levels
    WS_SPH_OPS,     // spherical FMM: dense translation operators
    WS_SPH_GRP,     // spherical FMM: boxes grouped by octant, M2L tiles
    WS_SPH_NEAR,    // spherical FMM: P2P sums per source in key order, P2P work units
    WS_SPH_M2LP,    // spherical FMM: M2L partial sums of the offset chunks
    WS_SPH_CELLS,   // spherical FMM: cell -> box table of one ternary level
    WS_NSLOTS
};

}  // namespace shp

namespace shp {
// Device copy of the background-field mesh plus its point locator: a quadtree
// over the (r, z) bounding square whose leaves list the triangles whose
// bounding boxes touch them (a leaf that touches none lists the triangle
// nearest to its centre, from which the reference walk continues).
struct FieldMesh {
    bool ready = false;
    int nnodes = 0, ntri = 0, nqnodes = 0, nqlist = 0;
    double* xy = nullptr;    // [nnodes][2]
    double* er = nullptr;    // [nnodes]
    double* ez = nullptr;    // [nnodes]
    int* tri = nullptr;      // [ntri][3]
    int* nbr = nullptr;      // [ntri][3]
    int2* qnode = nullptr;   // {first child | list offset, -1 | count}
    int* qlist = nullptr;
    double root[3] = {0, 0, 0};  // x0, y0, side
    double r0 = 0, E0 = 0;
};
}  // namespace shp

namespace shp {
__device__ __forceinline__ double rsqrt_f64(double x) {
    // v_rsq_f64 estimate refined by one third-order Newton step:
    // y <- y (1 + e/2 + 3e^2/8), e = 1 - x y^2.
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}
}  // namespace shp

struct shp_context {
    int device = 0;
    shp::FieldMesh mesh;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev2 = nullptr, ev3 = nullptr;  // update-kernel timing
    hipEvent_t fev[8] = {};                   // FMM pass timing
    unsigned long long* fmm_stats = nullptr;  // device counters of the last FMM call
    // FMM measurements: the last call's events and counters are folded into
    // fmm_last / fmm_acc once the stream has passed them (see shp_fmm_stats).
    bool fmm_pending = false, fmm_have_last = false, fmm_unrolled = false;
    int fmm_leaf_level = 0;
    double fmm_shell_madds[3] = {0, 0, 0};   // multiply-adds per M2L pair, shells 3/4/5
    double fmm_shell_madds0[3] = {0, 0, 0};  // the same, first M2L pass only
    double fmm_last[shp::kFmmStats] = {};
    double fmm_acc[shp::kFmmStats] = {};
    long fmm_acc_n = 0;
    // leaf-level choice of the last FMM call: rows (level, P2P pairs, far-band
    // box pairs, estimated ms), fmm_est_n of them
    int fmm_est_n = 0;
    int fmm_type = 2;  // FmmType of shp_fmm_set_type (SHP_FMM_*)
    double fmm_est[4 * 24] = {};
    bool timed = false, timed_update = false;
    // Barnes-Hut root-layout options (shp_bh_set_options; reference defaults)
    // and the last evaluation's tree summary (shp_bh_stats).
    shp_bh_options bh_opt = {0, 0, {1, 1, 1}, 2, 0, 0};
    bool bh_have_stats = false;
    unsigned long long* bh_ctr = nullptr;  // device counters in the workspace
    double bh_host_stats[11] = {};
    // Resident state of shp_state_upload_host / shp_state_run (own buffers,
    // apart from the workspace that the force evaluations use).
    double* res_f64 = nullptr;  // x y z vx vy vz q t_inj ax ay az
    int* res_i32 = nullptr;     // status, species
    float4* res_aos = nullptr;  // host AoS staging: position, velocity, acceleration, t_inj
    int res_n = 0, res_cap = 0;
    int* res_level = nullptr;   // ion levels of shp_state_run_substeps
    int res_level_cap = 0;
    // Outputs of the resident runs: each ion's background field at its last
    // kick ([3][res_cap] floats) and the queue of boundary-condition events
    // (res_cap of them: an ion leaves once per upload), kept while
    // res_events_on (inside shp_state_run*); drift_seq numbers the drifts
    // since the upload, which orders the events as the reference writes them.
    float* res_e2 = nullptr;
    unsigned int* res_ev = nullptr;
    unsigned long long* res_evkey = nullptr;
    int* res_evn = nullptr;
    bool res_events_on = false;
    long long drift_seq = 0;
    shp::StateWriter* writer = nullptr;
    // Spherical-harmonic FMM (shp_fmm_sph_host): device tables, built on first
    // use, and the last call's summary (shp_fmm_sph_stats).
    void* sph = nullptr;
    double sph_stats[12] = {};
    bool sph_have_stats = false;
    void* ws[shp::WS_NSLOTS] = {};
    size_t ws_bytes[shp::WS_NSLOTS] = {};
    int num_cus = 256;

    void* get(int slot, size_t bytes);  // grow-only; nullptr on failure (error set)
};

#define SHP_HIP_CHECK(expr)                                                           \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess) {                                                       \
            shp::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,          \
                           hipGetErrorString(e_));                                    \
            return SHP_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

#define SHP_REQUIRE(cond, ...)                                                        \
    do {                                                                              \
        if (!(cond)) {                                                                \
            shp::set_error(__VA_ARGS__);                                              \
            return SHP_ERR_INVALID;                                                   \
        }                                                                             \
    } while (0)

#define SHP_WS(ptr, type, h, slot, count)                                             \
    type* ptr = static_cast<type*>((h)->get((slot), sizeof(type) * (size_t)(count)));  \
    if (!ptr) return SHP_ERR_NOMEM
