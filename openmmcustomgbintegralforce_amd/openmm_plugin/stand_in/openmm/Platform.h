// OpenMM-shaped stand-in (used only when OpenMM is not installed): named platforms and
// their kernel-factory registry, as OpenMM::Platform::registerKernelFactory /
// createKernel.
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "openmm/KernelFactory.h"
#include "openmm/KernelImpl.h"
#include "openmm/OpenMMException.h"

namespace OpenMM {

class Platform {
public:
    explicit Platform(std::string name) : name_(std::move(name)) {}
    virtual ~Platform() = default;
    const std::string& getName() const { return name_; }

    // Takes ownership of the factory (shared between the names it serves).
    void registerKernelFactory(const std::string& kernel_name, KernelFactory* factory) {
        std::shared_ptr<KernelFactory> owned;
        for (auto& kv : factories_)
            if (kv.second.get() == factory) owned = kv.second;
        if (!owned) owned.reset(factory);
        factories_[kernel_name] = owned;
    }
    bool supportsKernels(const std::vector<std::string>& names) const {
        for (const auto& n : names)
            if (!factories_.count(n)) return false;
        return true;
    }
    KernelImpl* createKernel(const std::string& name, ContextImpl& context) const {
        auto it = factories_.find(name);
        if (it == factories_.end()) throw OpenMMException("Called createKernel() on a Platform which does not support the requested kernel");
        return it->second->createKernelImpl(name, *this, context);
    }

    static int getNumPlatforms() { return (int)registry().size(); }
    static Platform& getPlatform(int index) { return *registry().at(index); }
    static Platform& getPlatformByName(const std::string& name) {
        for (auto& p : registry())
            if (p->getName() == name) return *p;
        throw OpenMMException("There is no registered Platform called \"" + name + "\"");
    }
    static void registerPlatform(Platform* platform) { registry().emplace_back(platform); }

private:
    static std::vector<std::unique_ptr<Platform>>& registry() {
        static std::vector<std::unique_ptr<Platform>> platforms;
        return platforms;
    }
    std::string name_;
    std::map<std::string, std::shared_ptr<KernelFactory>> factories_;
};

}  // namespace OpenMM
