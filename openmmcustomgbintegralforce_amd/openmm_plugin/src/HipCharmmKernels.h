// OpenMM kernel classes of the HIP platform for CharmmGBSWForce and CharmmGBMVForce,
// implemented on the C ABI of include/gbi.h.  They replace the Reference platform's
// ReferenceCalcCharmmGB{SW,MV}ForceKernel (CharmmReferenceKernels.h) and the CUDA
// platform's CudaCalcCharmmGBMVForceKernel behind the interfaces of
// openmmapi/include/openmm/charmmKernels.h:49-129.
#pragma once

#include <string>
#include <vector>

#include "gbi.h"
#include "openmm/CharmmGBMVForce.h"
#include "openmm/CharmmGBSWForce.h"
#include "openmm/charmmKernels.h"

namespace gbi_openmm {

// Base of both kernels: one gbi context, positions in, forces added out.
class HipGBKernelBase {
public:
    virtual ~HipGBKernelBase();
    gbi_ctx* gbiContext() const { return ctx_; }

protected:
    // Device buffers of a GPU platform: positions are gathered from its posq and forces added
    // to its fixed-point buffer on the device (gbi_set_positions_posq, gbi_add_forces_fixed),
    // with no host round trip, as the CUDA platform's kernels read posq and add into its force
    // buffer (reduceIntegralForce.cu:61-77).
    struct PlatformBuffers {
        const void* posq = nullptr;
        int posq_double = 0;
        long long* force = nullptr;
        int padded_n = 0;
        const int* atom_slot = nullptr;
        void* stream = nullptr;
    };
    // Whether the Context keeps its state on the device, and its buffers.  The stand-in
    // ContextImpl reports its DeviceBuffers; a HIP OpenMM platform overrides this with its own.
    virtual bool platformBuffers(OpenMM::ContextImpl& context, PlatformBuffers& out);
    double run(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy);
    // Adds N x 3 forces (kJ/mol/nm) to the platform's force accumulator.  The OpenMM-shaped
    // stand-in ContextImpl keeps a host array; a HIP OpenMM platform overrides this and adds
    // gbi_get_device_view()'s int64 fixed-point forces to its own buffer on the device.
    virtual void addForces(OpenMM::ContextImpl& context, const std::vector<double>& forces);
    // Adds dE/d(global parameter) of the last evaluation to the Context's energy-parameter
    // derivatives (CharmmReferenceKernels.cpp:366-368).  The stand-in ContextImpl keeps a host
    // map; a HIP OpenMM platform overrides this with its own accumulator.
    virtual void addEnergyParameterDerivatives(OpenMM::ContextImpl& context, const std::vector<std::string>& names,
                                               const std::vector<double>& values);
    static void check(gbi_status status);
    void create(int n, const std::vector<double>& charge, const std::vector<double>& radius, const gbi_params& params);
    void update(const std::vector<double>& charge, const std::vector<double>& radius);

    gbi_ctx* ctx_ = nullptr;
    int n_ = 0;
    bool periodic_ = false;  // CutoffPeriodic: the Context's box goes to the engine every evaluation
    void* stream_ = nullptr; // the platform stream the engine works on
};

// CalcCharmmGBSWForceKernel on HIP (CharmmReferenceKernels.cpp:380-445 is the Reference one).
class HipCalcCharmmGBSWForceKernel : public OpenMM::CalcCharmmGBSWForceKernel, public HipGBKernelBase {
public:
    // honorDielectrics: use (1/soluteDielectric - 1/solventDielectric); the reference
    // hard-codes (1 - 1/80) (CharmmReferenceGBSW.cpp:177-178), which stays the default.
    HipCalcCharmmGBSWForceKernel(std::string name, const OpenMM::Platform& platform, bool honorDielectrics = false)
        : CalcCharmmGBSWForceKernel(name, platform), honor_(honorDielectrics) {}
    void initialize(const OpenMM::System& system, const OpenMM::CharmmGBSWForce& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const OpenMM::CharmmGBSWForce& force) override;

private:
    bool honor_;
};

// CalcCharmmGBMVForceKernel on HIP: the canonical GB energy (integrals of order 4 and 7,
// B = slope/(alpha0 (1/eta - I4) + alpha1 (1/(4 eta^4) - I7)^(1/4)) + shift, STILL energy)
// runs on the hand-written kernels; other expressions on kernels generated from them
// (gbi_create_custom).
class HipCalcCharmmGBMVForceKernel : public OpenMM::CalcCharmmGBMVForceKernel, public HipGBKernelBase {
public:
    HipCalcCharmmGBMVForceKernel(std::string name, const OpenMM::Platform& platform)
        : CalcCharmmGBMVForceKernel(name, platform) {}
    void initialize(const OpenMM::System& system, const OpenMM::CharmmGBMVForce& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const OpenMM::CharmmGBMVForce& force) override;

private:
    void createCustom(const OpenMM::CharmmGBMVForce& force);
    // the Context's global parameters, pushed to the engine when they change
    void pullGlobals(OpenMM::ContextImpl& context);
    bool custom_ = false;
    gbi_params params_{};                   // canonical: the engine's constants
    std::vector<std::string> globals_;      // global parameter names of the force
    std::vector<double> global_values_;     // their values in the engine
    std::vector<std::string> derivs_;       // energy parameter derivatives asked for
};

// Whether a CharmmGBMVForce holds exactly the canonical expressions (by expression text and
// integral names: the Born expression names I4 and I7, so the integral named I4 must be of
// order 4 and the one named I7 of order 7) and asks for no energy parameter derivatives.
bool isCanonicalGBMV(const OpenMM::CharmmGBMVForce& force);

// The canonical CharmmGBMVForce expressions recognised by HipCalcCharmmGBMVForceKernel.
extern const char* const kCanonicalBornValue;  // computed value "B"
extern const char* const kCanonicalSelfEnergy;
extern const char* const kCanonicalPairEnergy;

// gbi_params, charges and radii of a canonical CharmmGBMVForce; throws OpenMMException
// otherwise.
void gbmvParamsFromForce(const OpenMM::CharmmGBMVForce& force, gbi_params& params, std::vector<double>& charge,
                         std::vector<double>& radius);

}  // namespace gbi_openmm
