// Stand-in of OpenMM's tabulated-function classes (openmm/TabulatedFunction.h), used only when
// OpenMM's headers are not on the include path: Continuous1D/2D/3DFunction and
// Discrete1D/2D/3DFunction.  Same constructors and accessors as OpenMM's.
#pragma once

#include <vector>

namespace OpenMM {

class TabulatedFunction {
public:
    virtual ~TabulatedFunction() = default;
    bool getPeriodic() const { return periodic_; }

protected:
    bool periodic_ = false;
};

class Continuous1DFunction : public TabulatedFunction {
public:
    Continuous1DFunction(const std::vector<double>& values, double min, double max, bool periodic = false)
        : values_(values), min_(min), max_(max) {
        periodic_ = periodic;
    }
    void getFunctionParameters(std::vector<double>& values, double& min, double& max) const {
        values = values_;
        min = min_;
        max = max_;
    }
    void setFunctionParameters(const std::vector<double>& values, double min, double max) {
        values_ = values;
        min_ = min;
        max_ = max;
    }

private:
    std::vector<double> values_;
    double min_, max_;
};

class Discrete1DFunction : public TabulatedFunction {
public:
    explicit Discrete1DFunction(const std::vector<double>& values) : values_(values) {}
    void getFunctionParameters(std::vector<double>& values) const { values = values_; }
    void setFunctionParameters(const std::vector<double>& values) { values_ = values; }

private:
    std::vector<double> values_;
};

class Discrete2DFunction : public TabulatedFunction {
public:
    Discrete2DFunction(int xsize, int ysize, const std::vector<double>& values)
        : xsize_(xsize), ysize_(ysize), values_(values) {}
    void getFunctionParameters(int& xsize, int& ysize, std::vector<double>& values) const {
        xsize = xsize_;
        ysize = ysize_;
        values = values_;
    }

private:
    int xsize_, ysize_;
    std::vector<double> values_;
};

class Discrete3DFunction : public TabulatedFunction {
public:
    Discrete3DFunction(int xsize, int ysize, int zsize, const std::vector<double>& values)
        : xsize_(xsize), ysize_(ysize), zsize_(zsize), values_(values) {}
    void getFunctionParameters(int& xsize, int& ysize, int& zsize, std::vector<double>& values) const {
        xsize = xsize_;
        ysize = ysize_;
        zsize = zsize_;
        values = values_;
    }

private:
    int xsize_, ysize_, zsize_;
    std::vector<double> values_;
};

class Continuous2DFunction : public TabulatedFunction {
public:
    Continuous2DFunction(int xsize, int ysize, const std::vector<double>& values, double xmin, double xmax, double ymin,
                         double ymax, bool periodic = false)
        : xsize_(xsize), ysize_(ysize), values_(values), xmin_(xmin), xmax_(xmax), ymin_(ymin), ymax_(ymax) {
        periodic_ = periodic;
    }
    void getFunctionParameters(int& xsize, int& ysize, std::vector<double>& values, double& xmin, double& xmax,
                               double& ymin, double& ymax) const {
        xsize = xsize_;
        ysize = ysize_;
        values = values_;
        xmin = xmin_;
        xmax = xmax_;
        ymin = ymin_;
        ymax = ymax_;
    }

private:
    int xsize_, ysize_;
    std::vector<double> values_;
    double xmin_, xmax_, ymin_, ymax_;
};

class Continuous3DFunction : public TabulatedFunction {
public:
    Continuous3DFunction(int xsize, int ysize, int zsize, const std::vector<double>& values, double xmin, double xmax,
                         double ymin, double ymax, double zmin, double zmax, bool periodic = false)
        : xsize_(xsize), ysize_(ysize), zsize_(zsize), values_(values), xmin_(xmin), xmax_(xmax), ymin_(ymin),
          ymax_(ymax), zmin_(zmin), zmax_(zmax) {
        periodic_ = periodic;
    }
    void getFunctionParameters(int& xsize, int& ysize, int& zsize, std::vector<double>& values, double& xmin,
                               double& xmax, double& ymin, double& ymax, double& zmin, double& zmax) const {
        xsize = xsize_;
        ysize = ysize_;
        zsize = zsize_;
        values = values_;
        xmin = xmin_;
        xmax = xmax_;
        ymin = ymin_;
        ymax = ymax_;
        zmin = zmin_;
        zmax = zmax_;
    }

private:
    int xsize_, ysize_, zsize_;
    std::vector<double> values_;
    double xmin_, xmax_, ymin_, ymax_, zmin_, zmax_;
};

}  // namespace OpenMM
