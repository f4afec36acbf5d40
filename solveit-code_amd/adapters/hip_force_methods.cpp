// Reference-side adapters to the MI355X engine (see hip_force_methods.h).
#include "hip_force_methods.h"

#include <cstdio>
#include <cstdlib>
#include <string>

namespace hip_adapter {

shp_symmetry symmetryOf(const ParticleSystem& ps, const Grid* grid, int nsec, std::vector<int>& store) {
    shp_symmetry sym{SHP_SYM_NONE, nsec, 0, nullptr, {0.0, 0.0, 0.0}};
    if (ps.params.symType == SymmetryType::rotational) {
        sym.type = SHP_SYM_ROTATIONAL;
    } else if (ps.params.symType == SymmetryType::translational) {
        if (grid == nullptr) fail("translational symmetry needs the Grid (setGrid)");
        store = expandLattice(*grid);
        sym.type = SHP_SYM_TRANSLATIONAL;
        sym.nimg = (int)(store.size() / 3);
        sym.lattice = store.data();
        sym.domain[0] = ps.params.domainSize.x;
        sym.domain[1] = ps.params.domainSize.y;
        sym.domain[2] = ps.params.domainSize.z;
    }
    return sym;
}

void addInto(ChargedParticles& cp, int first, int last, const std::vector<double>& out) {
    for (int i = first; i < last; i++) {
        const double* a = &out[3 * (size_t)(i - first)];
        cp.acceleration[i].x += (float)a[0];
        cp.acceleration[i].y += (float)a[1];
        cp.acceleration[i].z += (float)a[2];
    }
}

void fail(const char* what) {
    printf("\nError: %s: %s\n", what, shp_last_error());
    exit(1);
}

shp_handle createHandle(int device) {
    shp_handle h = nullptr;
    if (shp_create(device, &h) != SHP_OK) fail("shp_create");
    return h;
}

void loadFieldMesh(const ParticleSystem& ps, const char* folderName, const char* fieldsName,
                   const char* trianglesName, shp_field_mesh* mesh) {
    const std::string fields = std::string(folderName) + fieldsName;
    const std::string triangles = std::string(folderName) + trianglesName;
    if (shp_field_mesh_load_csv(fields.c_str(), triangles.c_str(), ps.params.r0, ps.params.E0, mesh) != SHP_OK)
        fail("shp_field_mesh_load_csv");
}

static void requireSystem(const ParticleSystem* ps) {
    if (ps == nullptr) {
        printf("Particle system not initialized.\n");
        exit(1);
    }
}

}  // namespace hip_adapter

using namespace hip_adapter;

// ------------------------------------------------------------------ direct ----

HipDirectForceCalculation::HipDirectForceCalculation(ParticleSystem* particleSystem, int device, int precision)
    : ForceCalculationMethod{particleSystem}, handle{createHandle(device)}, precision{precision} {
    setTimerName();
}

HipDirectForceCalculation::~HipDirectForceCalculation() { shp_destroy(handle); }

void HipDirectForceCalculation::updateAcceleration() {
    requireSystem(particleSystem);
    ChargedParticles& cp = particleSystem->chargedParticles;
    const shp_symmetry sym = symmetryOf(*particleSystem, grid, 4, lattice);
    out.assign(3 * (size_t)(lastParticle - firstParticle), 0.0);
    if (lastParticle > firstParticle &&
        shp_direct_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x, lastParticle,
                        firstParticle, lastParticle, firstParticle, lastParticle, precision, &sym,
                        out.data()) != SHP_OK)
        fail("shp_direct_host");
    addInto(cp, firstParticle, lastParticle, out);
}

void HipDirectForceCalculation::setSaveName() { saveName = "_directHIP.bin"; }
void HipDirectForceCalculation::setTimerName() { timerName = "Direct Method (MI355X)"; }

// -------------------------------------------------------------- Barnes-Hut ----

HipBarnesHutForceCalculation::HipBarnesHutForceCalculation(ParticleSystem* particleSystem, int device)
    : ForceCalculationMethod{particleSystem}, handle{createHandle(device)} {
    setTimerName();
}

HipBarnesHutForceCalculation::~HipBarnesHutForceCalculation() { shp_destroy(handle); }

void HipBarnesHutForceCalculation::applyOptions() {
    if (shp_bh_set_options(handle, &options) != SHP_OK) fail("shp_bh_set_options");
}

void HipBarnesHutForceCalculation::setBoxDimVals(vec3<int> boxDimVals) {
    options.force_box_dims = 1;
    options.box_dims[0] = boxDimVals.x;
    options.box_dims[1] = boxDimVals.y;
    options.box_dims[2] = boxDimVals.z;
    applyOptions();
}

void HipBarnesHutForceCalculation::setDomainToggle(DomainToggle toggle) {
    options.domain_toggle = toggle == DomainToggle::cubicCells ? 0 : 1;
    applyOptions();
}

void HipBarnesHutForceCalculation::setSubcellBlockSize(int subcellBlockSize) {
    options.subcell_block = subcellBlockSize;
    applyOptions();
}

void HipBarnesHutForceCalculation::setTheta(float tempTheta) { theta = tempTheta; }

void HipBarnesHutForceCalculation::setOutlierToggle(int toggle) {
    options.outlier_toggle = toggle;
    applyOptions();
}

void HipBarnesHutForceCalculation::updateAcceleration() {
    requireSystem(particleSystem);
    ChargedParticles& cp = particleSystem->chargedParticles;
    const shp_symmetry sym = symmetryOf(*particleSystem, grid, 4, lattice);
    out.assign(3 * (size_t)(lastParticle - firstParticle), 0.0);
    if (lastParticle > firstParticle &&
        shp_bh_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x, lastParticle,
                    firstParticle, lastParticle, theta, &sym, out.data()) != SHP_OK)
        fail("shp_bh_host");
    addInto(cp, firstParticle, lastParticle, out);
}

void HipBarnesHutForceCalculation::setSaveName() { saveName = "_barnesHutHIP.bin"; }
void HipBarnesHutForceCalculation::setTimerName() { timerName = "Barnes Hut Method (MI355X)"; }

// --------------------------------------------------------------------- FMM ----

HipMultipoleMethodForceCalculation::HipMultipoleMethodForceCalculation(ParticleSystem* particleSystem,
                                                                       int device, int order, int ring,
                                                                       int leafSize)
    : ForceCalculationMethod{particleSystem},
      handle{createHandle(device)},
      order{order},
      ring{ring},
      leafSize{leafSize} {
    setTimerName();
}

HipMultipoleMethodForceCalculation::~HipMultipoleMethodForceCalculation() { shp_destroy(handle); }

void HipMultipoleMethodForceCalculation::setFmmType(FmmType fType) {
    const int t = fType == FmmType::treecode           ? SHP_FMM_TREECODE
                  : fType == FmmType::treecodeOutliers ? SHP_FMM_TREECODE_OUTLIERS
                  : fType == FmmType::fmmOutliers      ? SHP_FMM_FMM_OUTLIERS
                                                       : SHP_FMM_FMM;
    if (shp_fmm_set_type(handle, t) != SHP_OK) fail("shp_fmm_set_type");
}

void HipMultipoleMethodForceCalculation::updateAcceleration() {
    requireSystem(particleSystem);
    ChargedParticles& cp = particleSystem->chargedParticles;
    const shp_symmetry sym = symmetryOf(*particleSystem, grid, nsec, lattice);
    out.assign(3 * (size_t)(lastParticle - firstParticle), 0.0);
    if (lastParticle > firstParticle) {
        if (referenceOperators) {
            // updateAccelerationTranslational reaches the 27 cells at shifts
            // {-1, 0, 1}^3 whatever numSymmetryLevels (>= 1) the Grid expands to
            shp_symmetry rs = sym;
            int lat27[81];
            if (sym.type == SHP_SYM_TRANSLATIONAL) {
                for (int k = 0; k < 27; k++) {
                    lat27[3 * k] = k / 9 - 1;
                    lat27[3 * k + 1] = (k / 3) % 3 - 1;
                    lat27[3 * k + 2] = k % 3 - 1;
                }
                rs.nimg = 27;
                rs.lattice = lat27;
            }
            if (shp_fmm_sph_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x,
                                 lastParticle, firstParticle, lastParticle, order, forcedMaxLevel, &rs,
                                 out.data()) != SHP_OK)
                fail("shp_fmm_sph_host");
        } else if (shp_fmm_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x,
                                lastParticle, firstParticle, lastParticle, order, leafLevel, leafSize, ring,
                                &sym, out.data()) != SHP_OK) {
            fail("shp_fmm_host");
        }
    }
    addInto(cp, firstParticle, lastParticle, out);
}

void HipMultipoleMethodForceCalculation::setSaveName() { saveName = "_fmmHIP.bin"; }
void HipMultipoleMethodForceCalculation::setTimerName() { timerName = "Fast Multipole Method (MI355X)"; }

// --------------------------------------------------------------- integrator ----

// The run's symmetry, boundary conditions and background field in the engine's options.
static shp_step_opts runOptions(ParticleSystem& particleSystem, const shp_step_opts& opts, const Grid* grid,
                                std::vector<int>& lattice, shp_handle handle, const shp_field_mesh* mesh) {
    shp_step_opts o = opts;
    o.sym = symmetryOf(particleSystem, grid, opts.sym.nsec > 0 ? opts.sym.nsec : 4, lattice);
    const BoundaryConditionParameters& bc = particleSystem.params.boundaryConditions;
    const bool space = bc.boundaryConditionType == BoundaryConditionOptions::space ||
                       bc.boundaryConditionType == BoundaryConditionOptions::space_and_time;
    const bool time = bc.boundaryConditionType == BoundaryConditionOptions::time ||
                      bc.boundaryConditionType == BoundaryConditionOptions::space_and_time;
    o.bc_space[0] = space ? bc.spatialBoundaryConditions.x : 0.0;
    o.bc_space[1] = space ? bc.spatialBoundaryConditions.y : 0.0;
    o.bc_space[2] = space ? bc.spatialBoundaryConditions.z : 0.0;
    o.bc_space[3] = space ? bc.spatialBoundaryConditions.w : 0.0;
    o.bc_time = time ? bc.temporalBoundaryCondition : 0.0;
    o.use_efield = particleSystem.interpScheme.bgField == BackgroundEfield::singleEmitter;
    if (o.use_efield) {
        if (mesh == nullptr) fail("the background field needs its mesh");
        if (shp_efield_setup(handle, mesh) != SHP_OK) fail("shp_efield_setup");
    }
    return o;
}

// What the reference's own writers read besides the state: the Coulomb and
// background fields of the step (Efield1 / Efield2, for saveData), and the
// step's boundary-condition events, written to the run's events file as
// writeToEventsFile writes them.
static void downloadOutputs(ParticleSystem& particleSystem, shp_handle handle) {
    ChargedParticles& cp = particleSystem.chargedParticles;
    if (cp.Efield1 != nullptr && cp.Efield2 != nullptr &&
        shp_state_fields_host(handle, &cp.Efield1[0].x, &cp.Efield2[0].x) != SHP_OK)
        fail("shp_state_fields_host");
    if (particleSystem.params.eventsFile != nullptr &&
        shp_state_events(handle, particleSystem.params.eventsFile, nullptr) != SHP_OK)
        fail("shp_state_events");
}

HipLeapfrogDKD::HipLeapfrogDKD(ParticleSystem& particleSystem, ForceCalculationMethod& forceMethod,
                               const shp_step_opts& opts, const shp_field_mesh* mesh, int device,
                               const Grid* grid)
    : Integrator{particleSystem, forceMethod}, handle{createHandle(device)} {
    energySaveName = "_LeapfrogDKD_HIP.bin";
    this->opts = runOptions(particleSystem, opts, grid, lattice, handle, mesh);
}

HipLeapfrogDKD::~HipLeapfrogDKD() { shp_destroy(handle); }

void HipLeapfrogDKD::runSimulation() {
    intervalCount = 0;
    prevTime = 0.0f;
    const float timeStep = particleSystem.params.timeStep;
    initializeSimulation();
    ChargedParticles& cp = particleSystem.chargedParticles;
    const int n = cp.maxNumParticles;
    if (shp_state_upload_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x,
                              cp.timeInjected, n) != SHP_OK)
        fail("shp_state_upload_host");
    for (int i = 1; i <= particleSystem.params.numIters; i++) {
        iterTimer.changeCount(i);
        tempPosTime = timeStep * (i - 1);
        tempVelTime = timeStep * (i - 1);
        if (shp_state_run(handle, &opts, (double)tempPosTime, (double)timeStep, 1) != SHP_OK)
            fail("shp_state_run");
        if (shp_state_download_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x) !=
            SHP_OK)
            fail("shp_state_download_host");
        downloadOutputs(particleSystem, handle);
        // ions injected by the end of the step (findMostRecentInjected, particles.cpp:1328-1359)
        setLastChargedParticle(cp.findMostRecentInjected(timeStep * i) + 1);
        currentTime = timeStep * i;
        finishIter(i);
    }
    finishFullSim();
}

// ---------------------------------------------------------- multi-rate ----

HipLeapfrogDKDSubsteps::HipLeapfrogDKDSubsteps(ParticleSystem& particleSystem, ForceCalculationMethod& forceMethod,
                                               const shp_step_opts& opts, const shp_field_mesh* mesh, int device,
                                               const Grid* grid, const shp_step_opts* substepOpts)
    : LeapfrogDKDSubsteps{particleSystem, forceMethod}, handle{createHandle(device)} {
    energySaveName = "_LeapfrogDKDSubsteps_HIP.bin";
    this->opts = runOptions(particleSystem, opts, grid, lattice, handle, mesh);
    subOpts = this->opts;
    if (substepOpts != nullptr) {
        subOpts.method = substepOpts->method;
        subOpts.precision = substepOpts->precision;
        subOpts.fmm_order = substepOpts->fmm_order;
        subOpts.leaf_size = substepOpts->leaf_size;
        subOpts.theta = substepOpts->theta;
        subOpts.fmm_ring = substepOpts->fmm_ring;
        subOpts.leaf_level = substepOpts->leaf_level;
    }
}

HipLeapfrogDKDSubsteps::~HipLeapfrogDKDSubsteps() { shp_destroy(handle); }

void HipLeapfrogDKDSubsteps::setSubstepTimes(float* substepTimes, BoundaryConditionParameters* bc, int size) {
    LeapfrogDKDSubsteps::setSubstepTimes(substepTimes, bc, size);
    float min_dt = substepTimes[0];
    for (int i = 1; i < size; i++) min_dt = substepTimes[i] < min_dt ? substepTimes[i] : min_dt;
    float temp_dt = particleSystem.params.timeStep;
    levelsM = 0;
    while (temp_dt > min_dt) {
        temp_dt /= 3.0;
        levelsM++;
    }
}

void HipLeapfrogDKDSubsteps::runSimulation() {
    intervalCount = 0;
    prevTime = 0.0f;
    const float timeStep = particleSystem.params.timeStep;
    if (substepBoundaryConditions == nullptr) {
        setDefaultSubstepTimes();
        levelsM = 0;
    }
    std::vector<shp_substep_level> table((size_t)levelsM + 1);
    for (int l = 0; l <= levelsM; l++) {
        const BoundaryConditionParameters& b = substepBoundaryConditions[l].bc;
        table[l].bc_type = (int)b.boundaryConditionType;
        table[l].bc_space[0] = b.spatialBoundaryConditions.x;
        table[l].bc_space[1] = b.spatialBoundaryConditions.y;
        table[l].bc_space[2] = b.spatialBoundaryConditions.z;
        table[l].bc_space[3] = b.spatialBoundaryConditions.w;
        table[l].bc_time = b.temporalBoundaryCondition;
    }
    initializeSimulation();
    ChargedParticles& cp = particleSystem.chargedParticles;
    const int n = cp.maxNumParticles;
    if (shp_state_upload_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x,
                              cp.timeInjected, n) != SHP_OK)
        fail("shp_state_upload_host");
    for (int i = 1; i <= particleSystem.params.numIters; i++) {
        iterTimer.changeCount(i);
        tempPosTime = timeStep * (i - 1);
        tempVelTime = timeStep * (i - 1);
        if (shp_state_run_substeps(handle, &opts, &subOpts, table.data(), levelsM, i == 1, (double)tempPosTime,
                                   (double)timeStep, 1) != SHP_OK)
            fail("shp_state_run_substeps");
        if (shp_state_download_host(handle, &cp.position[0].x, &cp.velocity[0].x, &cp.acceleration[0].x) !=
            SHP_OK)
            fail("shp_state_download_host");
        downloadOutputs(particleSystem, handle);
        setLastChargedParticle(cp.findMostRecentInjected(timeStep * i) + 1);
        currentTime = timeStep * i;
        finishIter(i);
    }
    finishFullSim();
}
