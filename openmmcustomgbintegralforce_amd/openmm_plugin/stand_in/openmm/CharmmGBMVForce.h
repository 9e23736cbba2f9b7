// Stand-in of the plugin API class CharmmGBMVForce (used only when the reference plugin's
// openmmapi headers are not on the include path).  Same methods as
// openmmapi/include/openmm/CharmmGBMVForce.h (per-particle and global parameters, GB
// integrals with integer/real parameters, computed values, energy terms, exclusions,
// GBIntegralType, tabulated functions); ForceImpl creation is not part of the stand-in.
#pragma once

#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "openmm/Force.h"
#include "openmm/OpenMMException.h"
#include "openmm/TabulatedFunction.h"

namespace OpenMM {

class CharmmGBMVForce : public Force {
public:
    enum NonbondedMethod { NoCutoff = 0, CutoffNonPeriodic = 1, CutoffPeriodic = 2 };
    enum ComputationType { SingleParticle = 0, ParticlePair = 1, ParticlePairNoExclusions = 2 };
    enum GBIntegralType { GBSWIntegral = 0, GBMVIntegralTypeI = 1, GBMVIntegralTypeII = 2, CustomIntegral = 3 };

    CharmmGBMVForce() = default;
    int getNumParticles() const { return (int)particles_.size(); }
    int getNumExclusions() const { return (int)exclusions_.size(); }
    int getNumPerParticleParameters() const { return (int)param_names_.size(); }
    int getNumGlobalParameters() const { return (int)globals_.size(); }
    int getNumEnergyParameterDerivatives() const { return (int)derivs_.size(); }
    int getNumComputedValues() const { return (int)values_.size(); }
    int getNumGBIntegrals() const { return (int)integrals_.size(); }
    int getNumEnergyTerms() const { return (int)energy_.size(); }
    NonbondedMethod getNonbondedMethod() const { return method_; }
    void setNonbondedMethod(NonbondedMethod method) { method_ = method; }
    double getCutoffDistance() const { return cutoff_; }
    void setCutoffDistance(double distance) { cutoff_ = distance; }

    int addPerParticleParameter(const std::string& name) {
        param_names_.push_back(name);
        return (int)param_names_.size() - 1;
    }
    const std::string& getPerParticleParameterName(int index) const { return param_names_.at(index); }
    void setPerParticleParameterName(int index, const std::string& name) { param_names_.at(index) = name; }
    int addGlobalParameter(const std::string& name, double defaultValue) {
        globals_.push_back({name, defaultValue});
        return (int)globals_.size() - 1;
    }
    const std::string& getGlobalParameterName(int index) const { return globals_.at(index).first; }
    void setGlobalParameterName(int index, const std::string& name) { globals_.at(index).first = name; }
    double getGlobalParameterDefaultValue(int index) const { return globals_.at(index).second; }
    void setGlobalParameterDefaultValue(int index, double defaultValue) { globals_.at(index).second = defaultValue; }
    void addEnergyParameterDerivative(const std::string& name) { derivs_.push_back(name); }
    const std::string& getEnergyParameterDerivativeName(int index) const { return derivs_.at(index); }

    int addParticle(const std::vector<double>& parameters = std::vector<double>()) {
        particles_.push_back(parameters);
        return (int)particles_.size() - 1;
    }
    void getParticleParameters(int index, std::vector<double>& parameters) const { parameters = particles_.at(index); }
    void setParticleParameters(int index, const std::vector<double>& parameters) { particles_.at(index) = parameters; }

    int addComputedValue(const std::string& name, const std::string& expression, ComputationType type) {
        values_.push_back({name, expression, type});
        return (int)values_.size() - 1;
    }
    void getComputedValueParameters(int index, std::string& name, std::string& expression, ComputationType& type) const {
        const Term& t = values_.at(index);
        name = t.name;
        expression = t.expression;
        type = t.type;
    }
    void setComputedValueParameters(int index, const std::string& name, const std::string& expression,
                                    ComputationType type) {
        values_.at(index) = {name, expression, type};
    }
    int addGBIntegral(const std::string& name, const std::vector<int>& parametersInt = std::vector<int>(),
                      const std::vector<double>& parametersReal = std::vector<double>()) {
        integrals_.push_back({name, parametersInt, parametersReal});
        return (int)integrals_.size() - 1;
    }
    void getGBIntegralParameters(int index, std::string& name, std::vector<int>& parametersInt,
                                 std::vector<double>& parametersReal) const {
        const Integral& g = integrals_.at(index);
        name = g.name;
        parametersInt = g.ints;
        parametersReal = g.reals;
    }
    void getGBIntegralParameters(int index, std::string& name, std::vector<int>& parametersInt) const {
        std::vector<double> reals;
        getGBIntegralParameters(index, name, parametersInt, reals);
    }
    void getGBIntegralParameters(int index, std::string& name) const { name = integrals_.at(index).name; }
    void setGBIntegralParameters(int index, const std::string& name, std::vector<int> parametersInt,
                                 const std::vector<double> parametersReal) {
        integrals_.at(index) = {name, parametersInt, parametersReal};
    }
    void setGBIntegralType(GBIntegralType type) { integral_type_ = type; }
    GBIntegralType getGBIntegralType() const { return integral_type_; }
    int addEnergyTerm(const std::string& expression, ComputationType type) {
        energy_.push_back({"", expression, type});
        return (int)energy_.size() - 1;
    }
    void getEnergyTermParameters(int index, std::string& expression, ComputationType& type) const {
        expression = energy_.at(index).expression;
        type = energy_.at(index).type;
    }
    void setEnergyTermParameters(int index, const std::string& expression, ComputationType type) {
        energy_.at(index) = {"", expression, type};
    }
    int addExclusion(int particle1, int particle2) {
        exclusions_.push_back({particle1, particle2});
        return (int)exclusions_.size() - 1;
    }
    void getExclusionParticles(int index, int& particle1, int& particle2) const {
        particle1 = exclusions_.at(index).first;
        particle2 = exclusions_.at(index).second;
    }
    bool usesPeriodicBoundaryConditions() const { return method_ == CutoffPeriodic; }

    // tabulated functions (the force takes ownership, as OpenMM's forces do)
    int addTabulatedFunction(const std::string& name, TabulatedFunction* function) {
        functions_.push_back({name, std::shared_ptr<TabulatedFunction>(function)});
        return (int)functions_.size() - 1;
    }
    int getNumTabulatedFunctions() const { return (int)functions_.size(); }
    int getNumFunctions() const { return getNumTabulatedFunctions(); }
    const TabulatedFunction& getTabulatedFunction(int index) const { return *functions_.at(index).second; }
    TabulatedFunction& getTabulatedFunction(int index) { return *functions_.at(index).second; }
    const std::string& getTabulatedFunctionName(int index) const { return functions_.at(index).first; }
    int addFunction(const std::string& name, const std::vector<double>& values, double min, double max) {
        return addTabulatedFunction(name, new Continuous1DFunction(values, min, max));
    }

private:
    struct Term {
        std::string name, expression;
        ComputationType type;
    };
    struct Integral {
        std::string name;
        std::vector<int> ints;
        std::vector<double> reals;
    };
    std::vector<std::vector<double>> particles_;
    std::vector<std::string> param_names_, derivs_;
    std::vector<std::pair<std::string, double>> globals_;
    std::vector<Term> values_, energy_;
    std::vector<Integral> integrals_;
    std::vector<std::pair<int, int>> exclusions_;
    std::vector<std::pair<std::string, std::shared_ptr<TabulatedFunction>>> functions_;
    NonbondedMethod method_ = NoCutoff;
    GBIntegralType integral_type_ = GBSWIntegral;
    double cutoff_ = 1.0;
};

// The reference builds its family on the same class (CustomGBIntegralForce usage).
using CustomGBIntegralForce = CharmmGBMVForce;

}  // namespace OpenMM
