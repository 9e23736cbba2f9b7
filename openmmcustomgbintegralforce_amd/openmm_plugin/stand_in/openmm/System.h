// OpenMM-shaped stand-in (used only when OpenMM is not installed): particle masses only.
#pragma once

#include <vector>

namespace OpenMM {

class System {
public:
    int addParticle(double mass) {
        masses_.push_back(mass);
        return (int)masses_.size() - 1;
    }
    int getNumParticles() const { return (int)masses_.size(); }
    double getParticleMass(int index) const { return masses_.at(index); }

private:
    std::vector<double> masses_;
};

}  // namespace OpenMM
