// OpenMM-shaped stand-in (used only when OpenMM is not installed): a 3-vector of doubles
// with the member access the HIP kernels need.
#pragma once

namespace OpenMM {

class Vec3 {
public:
    Vec3() : d_{0.0, 0.0, 0.0} {}
    Vec3(double x, double y, double z) : d_{x, y, z} {}
    double operator[](int i) const { return d_[i]; }
    double& operator[](int i) { return d_[i]; }
    Vec3& operator+=(const Vec3& o) {
        for (int i = 0; i < 3; ++i) d_[i] += o.d_[i];
        return *this;
    }

private:
    double d_[3];
};

}  // namespace OpenMM
