// OpenMM-shaped stand-in (used only when OpenMM is not installed).
#pragma once

#include <string>

namespace OpenMM {

class ContextImpl;
class KernelImpl;
class Platform;

class KernelFactory {
public:
    virtual ~KernelFactory() = default;
    virtual KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const = 0;
};

}  // namespace OpenMM
