// OpenMM-shaped stand-in (used only when OpenMM is not installed): the positions of a
// Context, its global parameters, the host force accumulator a kernel adds into (the role of
// ReferencePlatform::PlatformData's force array) and the energy-parameter derivatives a kernel
// adds into (the map the Reference platform reaches through extractEnergyParameterDerivatives).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "openmm/OpenMMException.h"
#include "openmm/Vec3.h"

namespace OpenMM {

// The device buffers of a GPU platform's Context (OpenMM's CUDA and HIP platforms keep posq
// and an int64 fixed-point force buffer on the device, atoms in the platform's own order).
struct DeviceBuffers {
    const void* posq = nullptr;   // (x, y, z, q) per slot, float4 or double4
    int posq_double = 0;
    long long* force = nullptr;   // force[d * padded_n + slot], scale 2^32
    int padded_n = 0;
    const int* atom_slot = nullptr;  // slot of each atom (device), nullptr: identity
    void* stream = nullptr;       // the platform's stream (hipStream_t), nullptr: the null stream
};

class ContextImpl {
public:
    explicit ContextImpl(int n_particles) : positions_(n_particles), forces_(n_particles) {}
    void getPositions(std::vector<Vec3>& positions) const { positions = positions_; }
    void setPositions(const std::vector<Vec3>& positions) { positions_ = positions; }
    void getPeriodicBoxVectors(Vec3& a, Vec3& b, Vec3& c) const {
        a = box_[0];
        b = box_[1];
        c = box_[2];
    }
    void setPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c) {
        box_[0] = a;
        box_[1] = b;
        box_[2] = c;
    }
    // Global parameters: a Context holds one value per global parameter of its forces,
    // starting at the forces' defaults (ContextImpl's constructor in OpenMM).
    double getParameter(const std::string& name) const {
        auto it = parameters_.find(name);
        if (it == parameters_.end()) throw OpenMMException("Called getParameter() with invalid parameter name");
        return it->second;
    }
    void setParameter(const std::string& name, double value) { parameters_[name] = value; }
    std::map<std::string, double>& energyParameterDerivatives() { return param_derivs_; }
    std::vector<Vec3>& getForces() { return forces_; }
    // set when the Context's state lives on the device (a GPU platform), empty otherwise
    DeviceBuffers& deviceBuffers() { return device_; }
    void systemChanged() {}

private:
    DeviceBuffers device_;
    std::vector<Vec3> positions_, forces_;
    std::map<std::string, double> parameters_, param_derivs_;
    Vec3 box_[3] = {Vec3(2.0, 0.0, 0.0), Vec3(0.0, 2.0, 0.0), Vec3(0.0, 0.0, 2.0)};
};

}  // namespace OpenMM
