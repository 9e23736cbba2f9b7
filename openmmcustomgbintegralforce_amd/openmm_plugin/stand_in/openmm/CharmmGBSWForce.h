// Stand-in of the plugin API class CharmmGBSWForce (used only when the reference plugin's
// openmmapi headers are not on the include path).  Same methods as
// openmmapi/include/openmm/CharmmGBSWForce.h:58-203; ForceImpl creation is OpenMM's and
// is not part of the stand-in.
#pragma once

#include <vector>

#include "openmm/Force.h"
#include "openmm/OpenMMException.h"

namespace OpenMM {

class CharmmGBSWForce : public Force {
public:
    enum NonbondedMethod { NoCutoff = 0, CutoffNonPeriodic = 1, CutoffPeriodic = 2 };

    CharmmGBSWForce() = default;
    int getNumParticles() const { return (int)particles_.size(); }
    int addParticle(double charge, double radius, double scalingFactor) {
        particles_.push_back({charge, radius, scalingFactor});
        return (int)particles_.size() - 1;
    }
    void getParticleParameters(int index, double& charge, double& radius, double& scalingFactor) const {
        const Particle& p = at(index);
        charge = p.charge;
        radius = p.radius;
        scalingFactor = p.scaling;
    }
    void setParticleParameters(int index, double charge, double radius, double scalingFactor) {
        at(index);
        particles_[index] = {charge, radius, scalingFactor};
    }
    double getSolventDielectric() const { return solvent_; }
    void setSolventDielectric(double dielectric) { solvent_ = dielectric; }
    double getSoluteDielectric() const { return solute_; }
    void setSoluteDielectric(double dielectric) { solute_ = dielectric; }
    double getSurfaceAreaEnergy() const { return surface_; }
    void setSurfaceAreaEnergy(double energy) { surface_ = energy; }
    NonbondedMethod getNonbondedMethod() const { return method_; }
    void setNonbondedMethod(NonbondedMethod method) { method_ = method; }
    double getCutoffDistance() const { return cutoff_; }
    void setCutoffDistance(double distance) { cutoff_ = distance; }
    bool usesPeriodicBoundaryConditions() const { return method_ == CutoffPeriodic; }

private:
    struct Particle {
        double charge, radius, scaling;
    };
    const Particle& at(int index) const {
        if (index < 0 || index >= (int)particles_.size()) throw OpenMMException("CharmmGBSWForce: particle index out of range");
        return particles_[index];
    }
    std::vector<Particle> particles_;
    NonbondedMethod method_ = NoCutoff;
    double cutoff_ = 1.0, solvent_ = 78.3, solute_ = 1.0, surface_ = 2.25936;
};

}  // namespace OpenMM
