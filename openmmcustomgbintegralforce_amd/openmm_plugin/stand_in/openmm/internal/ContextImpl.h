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
    std::vector<Vec3>& getForces() { return forces_; }
    void systemChanged() {}

private:
    std::vector<Vec3> positions_, forces_;
    Vec3 box_[3] = {Vec3(2.0, 0.0, 0.0), Vec3(0.0, 2.0, 0.0), Vec3(0.0, 0.0, 2.0)};
};

}  // namespace OpenMM
