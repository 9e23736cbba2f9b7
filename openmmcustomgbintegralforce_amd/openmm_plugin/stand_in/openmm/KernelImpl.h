// OpenMM-shaped stand-in (used only when OpenMM is not installed).
#pragma once

#include <string>

namespace OpenMM {

class Platform;

class KernelImpl {
public:
    KernelImpl(std::string name, const Platform& platform) : name_(std::move(name)), platform_(&platform) {}
    virtual ~KernelImpl() = default;
    const std::string& getName() const { return name_; }
    const Platform& getPlatform() const { return *platform_; }

private:
    std::string name_;
    const Platform* platform_;
};

}  // namespace OpenMM
