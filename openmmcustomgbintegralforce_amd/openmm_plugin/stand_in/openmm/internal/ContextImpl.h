// OpenMM-shaped stand-in (used only when OpenMM is not installed): the positions of a
// Context and the host force accumulator a kernel adds into (the role of
// ReferencePlatform::PlatformData's force array).
#pragma once

#include <vector>

#include "openmm/Vec3.h"

namespace OpenMM {

class ContextImpl {
public:
    explicit ContextImpl(int n_particles) : positions_(n_particles), forces_(n_particles) {}
    void getPositions(std::vector<Vec3>& positions) const { positions = positions_; }
    void setPositions(const std::vector<Vec3>& positions) { positions_ = positions; }
    std::vector<Vec3>& getForces() { return forces_; }
    void systemChanged() {}

private:
    std::vector<Vec3> positions_, forces_;
};

}  // namespace OpenMM
