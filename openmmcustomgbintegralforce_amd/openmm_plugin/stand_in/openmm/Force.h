// OpenMM-shaped stand-in (used only when OpenMM is not installed): the force-group part
// of OpenMM::Force.  Creating a ForceImpl belongs to OpenMM's Context and is not part of
// this stand-in.
#pragma once

namespace OpenMM {

class Force {
public:
    virtual ~Force() = default;
    int getForceGroup() const { return group_; }
    void setForceGroup(int group) { group_ = group; }

private:
    int group_ = 0;
};

}  // namespace OpenMM
