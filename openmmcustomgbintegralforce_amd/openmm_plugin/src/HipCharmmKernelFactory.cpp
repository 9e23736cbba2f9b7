// Kernel factory and plugin entry points of the HIP CharmmGB kernels.  Mirrors
// CharmmReferenceKernelFactory.cpp:35-71 / CharmmCudaKernelFactory.cpp:35-69: the factory
// is registered on the platform named "HIP" for both kernel names; other names throw.
#include <string>

#include "HipCharmmKernels.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"

using namespace OpenMM;

namespace gbi_openmm {

class HipCharmmKernelFactory : public KernelFactory {
public:
    KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const override {
        (void)context;
        if (name == CalcCharmmGBMVForceKernel::Name()) return new HipCalcCharmmGBMVForceKernel(name, platform);
        if (name == CalcCharmmGBSWForceKernel::Name()) return new HipCalcCharmmGBSWForceKernel(name, platform);
        throw OpenMMException((std::string("Tried to create kernel with illegal kernel name '") + name + "'").c_str());
    }
};

}  // namespace gbi_openmm

extern "C" void registerPlatforms() {}

extern "C" void registerKernelFactories() {
    for (int i = 0; i < Platform::getNumPlatforms(); ++i) {
        Platform& platform = Platform::getPlatform(i);
        if (platform.getName() == "HIP") {
            auto* factory = new gbi_openmm::HipCharmmKernelFactory();
            platform.registerKernelFactory(CalcCharmmGBMVForceKernel::Name(), factory);
            platform.registerKernelFactory(CalcCharmmGBSWForceKernel::Name(), factory);
        }
    }
}

extern "C" void registerCharmmHipKernelFactories() { registerKernelFactories(); }
