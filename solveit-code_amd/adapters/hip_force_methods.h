// Reference-side adapters: subclasses of SOLVEiT's ForceCalculationMethod
// (force_calculation_methods.h:28-87) and Integrator (integration_schemes.h:21-87)
// that forward the driver's own particle arrays to the MI355X engine's C ABI
// (include/solveit_hip.h).  They are compiled with the reference's sources
// and headers and linked against libsolveit_hip.so; no reference file
// changes.  Errors follow the reference: printf + exit(1)
// (direct_force_method.cpp:351-354).
#ifndef HIP_FORCE_METHODS_H
#define HIP_FORCE_METHODS_H

#include <vector>

#include "force_calculation_methods.h"
#include "integration_schemes.h"
#include "hip_lattice.h"
#include "solveit_hip.h"

namespace hip_adapter {

// The shp_symmetry of the run (params.symType): none, rotational (nsec
// sectors; the reference's quadrants are 4) or translational (the expanded
// lattice of `grid` times params.domainSize).  `store` keeps the lattice.
shp_symmetry symmetryOf(const ParticleSystem& ps, const Grid* grid, int nsec, std::vector<int>& store);

// acceleration[i].xyz += out rows (float additions, as the reference adds).
void addInto(ChargedParticles& cp, int first, int last, const std::vector<double>& out);

[[noreturn]] void fail(const char* what);

shp_handle createHandle(int device);

// The run's field mesh from the files the driver hands to
// ParticleSystem::initializeBackgroundEField (particles.cpp:452-463): folder
// and file names joined as the reference joins them, r0 and E0 from
// params.  Release it with shp_field_mesh_free.
void loadFieldMesh(const ParticleSystem& ps, const char* folderName, const char* fieldsName,
                   const char* trianglesName, shp_field_mesh* mesh);

}  // namespace hip_adapter

// Replaces DirectForceCalculation (direct_force_method.cpp:347-428): the
// softened sum over [firstParticle, lastParticle), plain, rotational or
// translational as params.symType says.  precision: SHP_PREC_REF (the
// reference's float-input, double-sum arithmetic) or SHP_PREC_FP32C.
class HipDirectForceCalculation : public ForceCalculationMethod {
public:
    HipDirectForceCalculation(ParticleSystem* particleSystem, int device = 0, int precision = SHP_PREC_REF);
    ~HipDirectForceCalculation();
    void updateAcceleration() override;
    void increaseLevel() override {}
    void setSaveName() override;
    void setTimerName() override;

protected:
    void allocateThreadMemory() override {}
    void deallocateThreadMemory() override {}

private:
    shp_handle handle = nullptr;
    int precision;
    std::vector<double> out;
    std::vector<int> lattice;
};

// Replaces BarnesHutForceCalculation (barnes_hut_force_calculation.cpp:574-715)
// with its setters (force_calculation_methods.h:222-329).
class HipBarnesHutForceCalculation : public ForceCalculationMethod {
public:
    HipBarnesHutForceCalculation(ParticleSystem* particleSystem, int device = 0);
    ~HipBarnesHutForceCalculation();
    void updateAcceleration() override;
    void increaseLevel() override {}
    void setSaveName() override;
    void setTimerName() override;

    void setBoxDimVals(vec3<int> boxDimVals);
    void setDomainToggle(DomainToggle toggle);
    void setSubcellBlockSize(int subcellBlockSize);
    void setTheta(float tempTheta);
    void setOutlierToggle(int toggle);

protected:
    void allocateThreadMemory() override {}
    void deallocateThreadMemory() override {}

private:
    void applyOptions();
    shp_handle handle = nullptr;
    float theta = BHConstant::THETA;
    shp_bh_options options = {0, 0, {1, 1, 1}, 2, 0, 0};
    std::vector<double> out;
    std::vector<int> lattice;
};

// Replaces MultipoleMethodForceCalculation (multipole_method_force_calculation.cpp:148-182)
// and its rotational / translational siblings.  setNumExpansions sets the
// order p; setMaxLevel fixes the leaf level (0: the engine chooses it from
// the ion positions, at most leafSize ions per leaf on average as a start).
class HipMultipoleMethodForceCalculation : public ForceCalculationMethod {
public:
    HipMultipoleMethodForceCalculation(ParticleSystem* particleSystem, int device = 0, int order = 10,
                                       int ring = 2, int leafSize = 20);
    ~HipMultipoleMethodForceCalculation();
    void updateAcceleration() override;
    void increaseLevel() override {
        if (leafLevel > 0) leafLevel++;
    }
    void setSaveName() override;
    void setTimerName() override;

    void setNumExpansions(int p) { order = p; }
    void setMaxLevel(int maxLevel) {
        leafLevel = maxLevel;
        forcedMaxLevel = maxLevel;
    }
    void setNumSectors(int n) { nsec = n; }
    // The reference's FmmType (its constructor's fType, structures.h:283-288):
    // treecode or FMM, with or without outliers held out of the tree.
    void setFmmType(FmmType fType);
    // true: evaluate with the reference's own operators (spherical harmonics,
    // rotation-based translations, FP64; shp_fmm_sph_host) on the reference's
    // tree, with setMaxLevel as the reference's forcedMaxLevel; false (the
    // default): the engine's Cartesian expansion of the softened kernel.
    // Without symmetry or with the rotational one of 4 quadrants
    // (updateAccelerationRotational).
    void setReferenceOperators(bool on) { referenceOperators = on; }

protected:
    void allocateThreadMemory() override {}
    void deallocateThreadMemory() override {}

private:
    shp_handle handle = nullptr;
    int order, ring, leafSize, leafLevel = 0, nsec = 4;
    int forcedMaxLevel = -2;
    bool referenceOperators = false;
    std::vector<double> out;
    std::vector<int> lattice;
};

// LeapfrogDKD (integration_schemes.cpp:856-891) with the plume resident on the
// GPU: each iteration runs one DKD step on the device (force of the chosen
// method, background field, injection clamps, folding and boundary
// conditions), then writes positions, velocities, the Coulomb acceleration
// and the status back into the driver's arrays so that finishIter's output,
// energy and comparison bookkeeping see the step's result.  The background
// field needs the mesh of the run (interpolation_scheme.h:44-66), given as an
// shp_field_mesh.
class HipLeapfrogDKD : public Integrator {
public:
    HipLeapfrogDKD(ParticleSystem& particleSystem, ForceCalculationMethod& forceMethod,
                   const shp_step_opts& opts, const shp_field_mesh* mesh = nullptr, int device = 0,
                   const Grid* grid = nullptr);
    ~HipLeapfrogDKD();
    void runSimulation() override;

private:
    shp_handle handle = nullptr;
    shp_step_opts opts;
    std::vector<int> lattice;
};

// LeapfrogDKDSubsteps (integration_schemes.cpp:899-1529) with the plume
// resident on the GPU.  The level table is the reference's own: the base
// class's setSubstepTimes builds substepBoundaryConditions; the number of
// levels is counted as setSubstepTimes counts it (:1379-1385).  Each
// iteration runs on the device (shp_state_run_substeps) and writes the state
// back into the driver's arrays before finishIter, as HipLeapfrogDKD does.
// substepOpts: the force of the kicks below the top level (the reference's
// substepMethod); nullptr = opts.
class HipLeapfrogDKDSubsteps : public LeapfrogDKDSubsteps {
public:
    HipLeapfrogDKDSubsteps(ParticleSystem& particleSystem, ForceCalculationMethod& forceMethod,
                           const shp_step_opts& opts, const shp_field_mesh* mesh = nullptr, int device = 0,
                           const Grid* grid = nullptr, const shp_step_opts* substepOpts = nullptr);
    ~HipLeapfrogDKDSubsteps();
    void setSubstepTimes(float* substepTimes, BoundaryConditionParameters* bc, int size);
    void runSimulation() override;

private:
    shp_handle handle = nullptr;
    shp_step_opts opts, subOpts;
    std::vector<int> lattice;
    int levelsM = 0;
};

#endif  // HIP_FORCE_METHODS_H
