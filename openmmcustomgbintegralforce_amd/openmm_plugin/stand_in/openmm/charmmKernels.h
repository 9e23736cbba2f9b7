// Stand-in of the plugin's kernel interfaces (used only when the reference plugin's
// openmmapi headers are not on the include path): same classes, names and pure virtuals as
// openmmapi/include/openmm/charmmKernels.h:49-129.
#pragma once

#include <string>

#include "openmm/CharmmGBMVForce.h"
#include "openmm/CharmmGBSWForce.h"
#include "openmm/KernelImpl.h"
#include "openmm/System.h"
#include "openmm/internal/ContextImpl.h"

namespace OpenMM {

class CalcCharmmGBMVForceKernel : public KernelImpl {
public:
    enum NonbondedMethod { NoCutoff = 0, CutoffNonPeriodic = 1, CutoffPeriodic = 2 };
    static std::string Name() { return "CalcCharmmGBMVForce"; }
    CalcCharmmGBMVForceKernel(std::string name, const Platform& platform) : KernelImpl(name, platform) {}
    virtual void initialize(const System& system, const CharmmGBMVForce& force) = 0;
    virtual double execute(ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(ContextImpl& context, const CharmmGBMVForce& force) = 0;
};

class CalcCharmmGBSWForceKernel : public KernelImpl {
public:
    enum NonbondedMethod { NoCutoff = 0, CutoffNonPeriodic = 1, CutoffPeriodic = 2 };
    static std::string Name() { return "CalcCharmmGBSWForce"; }
    CalcCharmmGBSWForceKernel(std::string name, const Platform& platform) : KernelImpl(name, platform) {}
    virtual void initialize(const System& system, const CharmmGBSWForce& force) = 0;
    virtual double execute(ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(ContextImpl& context, const CharmmGBSWForce& force) = 0;
};

}  // namespace OpenMM
