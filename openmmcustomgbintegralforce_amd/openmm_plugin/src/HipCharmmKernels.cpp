// HIP kernels for CharmmGBSWForce / CharmmGBMVForce on the gbi C ABI.
#include "HipCharmmKernels.h"

#include <algorithm>
#include <cctype>

#include "openmm/OpenMMException.h"
#include "openmm/TabulatedFunction.h"

using namespace OpenMM;

namespace gbi_openmm {

const char* const kCanonicalBornValue = "slope*1/(alpha0*(1/eta-I4)+alpha1*(1/(4*eta^4)-I7)^(1/4))+shift";
const char* const kCanonicalSelfEnergy = "0.5*P*q^2/B";
const char* const kCanonicalPairEnergy = "P*q1*q2/sqrt(r^2+B1*B2*exp(-r^2/(4*B1*B2)))";

namespace {

std::string strip(const std::string& s) {
    std::string out;
    for (char c : s)
        if (!std::isspace((unsigned char)c)) out.push_back(c);
    return out;
}

// The force's per-particle parameters, N x n_params row-major.
std::vector<double> particleParameterArray(const CharmmGBMVForce& force) {
    const int n = force.getNumParticles(), np = force.getNumPerParticleParameters();
    std::vector<double> out((size_t)n * np);
    for (int i = 0; i < n; ++i) {
        std::vector<double> v;
        force.getParticleParameters(i, v);
        if ((int)v.size() < np) throw OpenMMException("CharmmGBMVForce: a particle lacks parameters");
        for (int k = 0; k < np; ++k) out[(size_t)i * np + k] = v[k];
    }
    return out;
}

int lebedevPoints(int rule) {
    static const int pts[] = {0, 6, 14, 26, 38, 50};
    if (rule < 1 || rule > 5) throw OpenMMException("CharmmGBMVForce: Lebedev rule must be 1..5 on the HIP platform");
    return pts[rule];
}

}  // namespace

HipGBKernelBase::~HipGBKernelBase() { gbi_destroy(ctx_); }

void HipGBKernelBase::check(gbi_status status) {
    if (status != GBI_OK) throw OpenMMException(gbi_last_error());
}

void HipGBKernelBase::create(int n, const std::vector<double>& charge, const std::vector<double>& radius,
                             const gbi_params& params) {
    gbi_destroy(ctx_);
    ctx_ = nullptr;
    n_ = n;
    periodic_ = params.nonbonded_method == GBI_CUTOFF_PERIODIC;
    check(gbi_create(n, charge.data(), radius.data(), &params, &ctx_));
}

void HipGBKernelBase::update(const std::vector<double>& charge, const std::vector<double>& radius) {
    // status GBI_ERR_PARTICLE_COUNT carries the reference's message (CharmmReferenceKernels.cpp:373-374)
    check(gbi_update_parameters(ctx_, (int32_t)charge.size(), charge.data(), radius.data()));
}

bool HipGBKernelBase::platformBuffers(ContextImpl& context, PlatformBuffers& out) {
#ifdef GBI_OPENMM_STAND_IN
    const DeviceBuffers& d = context.deviceBuffers();
    if (!d.posq) return false;
    out.posq = d.posq;
    out.posq_double = d.posq_double;
    out.force = d.force;
    out.padded_n = d.padded_n;
    out.atom_slot = d.atom_slot;
    out.stream = d.stream;
    return true;
#else
    (void)context;
    (void)out;
    return false;
#endif
}

double HipGBKernelBase::run(ContextImpl& context, bool includeForces, bool includeEnergy) {
    PlatformBuffers pb;
    if (platformBuffers(context, pb)) {
        if (pb.stream != stream_) {
            check(gbi_set_stream(ctx_, pb.stream));
            stream_ = pb.stream;
        }
        if (periodic_) {
            Vec3 a, b, c;
            context.getPeriodicBoxVectors(a, b, c);
            const double v[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
            check(gbi_set_periodic_box(ctx_, 0, v));
        }
        check(gbi_set_positions_posq(ctx_, pb.posq, pb.posq_double, pb.atom_slot));
        check(gbi_compute(ctx_, GBI_FORCES | GBI_ENERGY));
        check(gbi_synchronize(ctx_));  // device flags: errors, or buffer growth and a redo
        if (includeForces) check(gbi_add_forces_fixed(ctx_, reinterpret_cast<int64_t*>(pb.force), pb.padded_n,
                                                      pb.atom_slot));
        double energy = 0.0;
        if (includeEnergy) check(gbi_get_energy(ctx_, &energy));
        return energy;
    }
    std::vector<Vec3> pos;
    context.getPositions(pos);
    if ((int)pos.size() != n_) throw OpenMMException("HIP GB kernel: the Context's particle count changed");
    if (periodic_) {  // the box is read every evaluation (CharmmReferenceKernels.cpp:351, :430)
        Vec3 a, b, c;
        context.getPeriodicBoxVectors(a, b, c);
        const double v[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
        check(gbi_set_periodic_box(ctx_, 0, v));
    }
    std::vector<double> x(3 * (size_t)n_), f(3 * (size_t)n_, 0.0);
    for (int a = 0; a < n_; ++a)
        for (int d = 0; d < 3; ++d) x[3 * a + d] = pos[a][d];
    double energy = 0.0;
    check(gbi_execute(ctx_, x.data(), includeForces ? 1 : 0, includeEnergy ? 1 : 0, &energy,
                      includeForces ? f.data() : nullptr));
    if (includeForces) addForces(context, f);
    return includeEnergy ? energy : 0.0;
}

void HipGBKernelBase::addEnergyParameterDerivatives(ContextImpl& context, const std::vector<std::string>& names,
                                                    const std::vector<double>& values) {
#ifdef GBI_OPENMM_STAND_IN
    auto& acc = context.energyParameterDerivatives();
    for (size_t k = 0; k < names.size(); ++k) acc[names[k]] += values[k];
#else
    (void)context;
    (void)names;
    (void)values;
    throw OpenMMException("HIP GB kernel: the platform must override addEnergyParameterDerivatives()");
#endif
}

void HipGBKernelBase::addForces(ContextImpl& context, const std::vector<double>& forces) {
#ifdef GBI_OPENMM_STAND_IN
    std::vector<Vec3>& acc = context.getForces();
    for (int a = 0; a < n_; ++a) acc[a] += Vec3(forces[3 * a], forces[3 * a + 1], forces[3 * a + 2]);
#else
    (void)context;
    (void)forces;
    throw OpenMMException("HIP GB kernel: the platform must override addForces() to reach its force buffer");
#endif
}

// ---------------------------------------------------------------------------- GBSW

void HipCalcCharmmGBSWForceKernel::initialize(const System& system, const CharmmGBSWForce& force) {
    const int n = force.getNumParticles();
    if (system.getNumParticles() != n)  // CharmmGBSWForceImpl.cpp:47-48
        throw OpenMMException("CharmmGBSWForce must have exactly as many particles as the System it belongs to.");
    std::vector<double> q(n), r(n);
    for (int i = 0; i < n; ++i) {
        double s;
        force.getParticleParameters(i, q[i], r[i], s);
    }
    gbi_params p;
    check(gbi_params_default(GBI_MODEL_GBSW, 0, &p));
    if (honor_) p.prefactor = -138.935456 * (1.0 / force.getSoluteDielectric() - 1.0 / force.getSolventDielectric());
    p.nonbonded_method = (int32_t)force.getNonbondedMethod();
    p.cutoff = force.getCutoffDistance();
    create(n, q, r, p);
}

double HipCalcCharmmGBSWForceKernel::execute(ContextImpl& context, bool includeForces, bool includeEnergy) {
    return run(context, includeForces, includeEnergy);
}

void HipCalcCharmmGBSWForceKernel::copyParametersToContext(ContextImpl& context, const CharmmGBSWForce& force) {
    const int n = force.getNumParticles();
    std::vector<double> q(n), r(n);
    for (int i = 0; i < n; ++i) {
        double s;
        force.getParticleParameters(i, q[i], r[i], s);
    }
    update(q, r);
    context.systemChanged();
}

// ---------------------------------------------------------------------------- GBMV

void gbmvParamsFromForce(const CharmmGBMVForce& force, gbi_params& p, std::vector<double>& charge,
                         std::vector<double>& radius) {
    const CharmmGBMVForce::GBIntegralType type = force.getGBIntegralType();
    if (type == CharmmGBMVForce::CustomIntegral)  // CharmmReferenceKernels.cpp:129-133
        throw OpenMMException("CharmmGBMVForce: CustomIntegral (type 3) is not implemented by the reference either");
    int iq = -1, ir = -1;
    for (int k = 0; k < force.getNumPerParticleParameters(); ++k) {
        if (force.getPerParticleParameterName(k) == "q") iq = k;
        if (force.getPerParticleParameterName(k) == "radius") ir = k;
    }
    if (ir < 0)  // GBSWIntegral.cpp:79-80
        throw OpenMMException("GBSWIntegral: the per perticle parameter 'radius' must be defined");
    if (iq < 0) throw OpenMMException("CharmmGBMVForce: the canonical energy needs a per-particle parameter 'q'");
    std::vector<std::vector<int>> ints(force.getNumGBIntegrals());
    for (int k = 0; k < force.getNumGBIntegrals(); ++k) {
        std::string name;
        force.getGBIntegralParameters(k, name, ints[k]);
        if (ints[k].empty()) throw OpenMMException("GBSWIntegral: the order of integral must be given");
        if (ints[k][0] < 2) throw OpenMMException("GBSWIntegral: the order of integral must be greater or equal to 2");
    }
    if (!isCanonicalGBMV(force))
        throw OpenMMException("CharmmGBMVForce: only the canonical GB expressions run on the hand-written kernels");
    if (gbi_params_default(GBI_MODEL_GBMV, (int32_t)type, &p) != GBI_OK) throw OpenMMException(gbi_last_error());
    if (ints[0].size() > 1) p.n_lebedev = lebedevPoints(ints[0][1]);
    if (ints[0].size() > 2) p.n_radial = ints[0][2];
    p.orders[0] = ints[0][0];
    p.orders[1] = ints[1][0];
    p.nonbonded_method = (int32_t)force.getNonbondedMethod();
    p.cutoff = force.getCutoffDistance();
    bool seen[6] = {false, false, false, false, false, false};
    for (int k = 0; k < force.getNumGlobalParameters(); ++k) {
        const std::string& g = force.getGlobalParameterName(k);
        const double v = force.getGlobalParameterDefaultValue(k);
        if (g == "alpha0") p.alpha0 = v, seen[0] = true;
        else if (g == "alpha1") p.alpha1 = v, seen[1] = true;
        else if (g == "eta") p.eta = v, seen[2] = true;
        else if (g == "slope") p.slope = v, seen[3] = true;
        else if (g == "shift") p.shift = v, seen[4] = true;
        else if (g == "P") p.prefactor = v, seen[5] = true;
    }
    if (!std::all_of(seen, seen + 6, [](bool b) { return b; }))
        throw OpenMMException("CharmmGBMVForce: global parameters alpha0, alpha1, eta, slope, shift and P must be defined");
    const int n = force.getNumParticles();
    charge.resize(n);
    radius.resize(n);
    for (int i = 0; i < n; ++i) {
        std::vector<double> v;
        force.getParticleParameters(i, v);
        if ((int)v.size() <= std::max(iq, ir)) throw OpenMMException("CharmmGBMVForce: a particle lacks parameters");
        charge[i] = v[iq];
        radius[i] = v[ir];
    }
}

bool isCanonicalGBMV(const CharmmGBMVForce& force) {
    bool q = false;
    for (int k = 0; k < force.getNumPerParticleParameters(); ++k) q = q || force.getPerParticleParameterName(k) == "q";
    if (!q || force.getNumGBIntegrals() != 2 || force.getNumComputedValues() != 1 || force.getNumEnergyTerms() != 2)
        return false;
    if (force.getNumEnergyParameterDerivatives() > 0) return false;  // generated kernels compute them
    // the Born expression refers to the integrals by name
    for (int k = 0; k < 2; ++k) {
        std::string name;
        std::vector<int> o;
        force.getGBIntegralParameters(k, name, o);
        if (o.empty() || !((name == "I4" && o[0] == 4) || (name == "I7" && o[0] == 7))) return false;
    }
    std::string name, expr;
    CharmmGBMVForce::ComputationType t;
    force.getComputedValueParameters(0, name, expr, t);
    bool ok = name == "B" && strip(expr) == kCanonicalBornValue;
    force.getEnergyTermParameters(0, expr, t);
    ok = ok && strip(expr) == kCanonicalSelfEnergy && t == CharmmGBMVForce::SingleParticle;
    force.getEnergyTermParameters(1, expr, t);
    ok = ok && strip(expr) == kCanonicalPairEnergy &&
         (t == CharmmGBMVForce::ParticlePairNoExclusions || t == CharmmGBMVForce::ParticlePair);
    int seen = 0;
    for (int k = 0; k < force.getNumGlobalParameters(); ++k) {
        const std::string& g = force.getGlobalParameterName(k);
        seen += g == "alpha0" || g == "alpha1" || g == "eta" || g == "slope" || g == "shift" || g == "P";
    }
    return ok && seen == 6;
}

void HipCalcCharmmGBMVForceKernel::createCustom(const CharmmGBMVForce& force) {
    const CharmmGBMVForce::GBIntegralType type = force.getGBIntegralType();
    if (type == CharmmGBMVForce::CustomIntegral)  // CharmmReferenceKernels.cpp:129-133
        throw OpenMMException("CharmmGBMVForce: CustomIntegral (type 3) is not implemented by the reference either");
    std::vector<std::string> pnames, gnames, inames, vnames, vexprs, texprs;
    std::vector<double> gvals, pvals;
    std::vector<int32_t> ttypes;
    int radius = -1;
    for (int k = 0; k < force.getNumPerParticleParameters(); ++k) {
        pnames.push_back(force.getPerParticleParameterName(k));
        if (pnames.back() == "radius") radius = k;
    }
    if (radius < 0)  // GBSWIntegral.cpp:79-80
        throw OpenMMException("GBSWIntegral: the per perticle parameter 'radius' must be defined");
    gbi_params p;
    check(gbi_params_default(GBI_MODEL_GBMV, (int32_t)type, &p));
    for (int k = 0; k < force.getNumGBIntegrals(); ++k) {
        std::string name;
        std::vector<int> ints;
        force.getGBIntegralParameters(k, name, ints);
        if (ints.empty()) throw OpenMMException("GBSWIntegral: the order of integral must be given");
        if (ints[0] < 2) throw OpenMMException("GBSWIntegral: the order of integral must be greater or equal to 2");
        if (k == 0 && ints.size() > 1) p.n_lebedev = lebedevPoints(ints[1]);
        if (k == 0 && ints.size() > 2) p.n_radial = ints[2];
        if (k < 2) p.orders[k] = ints[0];
        inames.push_back(name);
    }
    for (int k = 0; k < force.getNumGlobalParameters(); ++k) {
        gnames.push_back(force.getGlobalParameterName(k));
        gvals.push_back(force.getGlobalParameterDefaultValue(k));
    }
    for (int k = 0; k < force.getNumComputedValues(); ++k) {
        std::string name, expr;
        CharmmGBMVForce::ComputationType t;
        force.getComputedValueParameters(k, name, expr, t);  // evaluated as SingleParticle (CharmmReferenceGBMV.cpp:173-176)
        vnames.push_back(name);
        vexprs.push_back(expr);
    }
    bool excl = false;
    for (int k = 0; k < force.getNumEnergyTerms(); ++k) {
        std::string expr;
        CharmmGBMVForce::ComputationType t;
        force.getEnergyTermParameters(k, expr, t);
        texprs.push_back(expr);
        ttypes.push_back((int32_t)t);
        excl = excl || t == CharmmGBMVForce::ParticlePair;
    }
    const int n = force.getNumParticles();
    const int np = (int)pnames.size();
    pvals = particleParameterArray(force);
    p.nonbonded_method = (int32_t)force.getNonbondedMethod();
    p.cutoff = force.getCutoffDistance();
    auto cstr = [](const std::vector<std::string>& v) {
        std::vector<const char*> out;
        for (const auto& x : v) out.push_back(x.c_str());
        return out;
    };
    // tabulated functions: OpenMM's Continuous1D/2D/3DFunction and Discrete1D/2D/3DFunction
    // (dims: the sizes of the tables of 2 or 3 arguments, values[x + xsize (y + ysize z)];
    // ranges_yz: the y and z ranges of the continuous ones)
    std::vector<std::string> fnames;
    std::vector<int32_t> fkinds, fsizes, fperiodic, fdims;
    std::vector<double> franges, fvalues, franges_yz;
    for (int k = 0; k < force.getNumTabulatedFunctions(); ++k) {
        const TabulatedFunction& fn = force.getTabulatedFunction(k);
        fnames.push_back(force.getTabulatedFunctionName(k));
        std::vector<double> v;
        double lo = 0.0, hi = 0.0, ylo = 0.0, yhi = 1.0, zlo = 0.0, zhi = 1.0;
        int nx = 0, ny = 1, nz = 1;
        if (const auto* c1 = dynamic_cast<const Continuous1DFunction*>(&fn)) {
            c1->getFunctionParameters(v, lo, hi);
            fkinds.push_back(0);
        } else if (const auto* d1 = dynamic_cast<const Discrete1DFunction*>(&fn)) {
            d1->getFunctionParameters(v);
            fkinds.push_back(1);
        } else if (const auto* d2 = dynamic_cast<const Discrete2DFunction*>(&fn)) {
            d2->getFunctionParameters(nx, ny, v);
            fkinds.push_back(2);
        } else if (const auto* d3 = dynamic_cast<const Discrete3DFunction*>(&fn)) {
            d3->getFunctionParameters(nx, ny, nz, v);
            fkinds.push_back(3);
        } else if (const auto* c2 = dynamic_cast<const Continuous2DFunction*>(&fn)) {
            c2->getFunctionParameters(nx, ny, v, lo, hi, ylo, yhi);
            fkinds.push_back(4);
        } else if (const auto* c3 = dynamic_cast<const Continuous3DFunction*>(&fn)) {
            c3->getFunctionParameters(nx, ny, nz, v, lo, hi, ylo, yhi, zlo, zhi);
            fkinds.push_back(5);
        } else {
            throw OpenMMException("CharmmGBMVForce: tabulated function '" + fnames.back() + "': unknown kind");
        }
        franges_yz.push_back(ylo);
        franges_yz.push_back(yhi);
        franges_yz.push_back(zlo);
        franges_yz.push_back(zhi);
        fdims.push_back(nx);
        fdims.push_back(ny);
        fdims.push_back(nz);
        fsizes.push_back((int32_t)v.size());
        fperiodic.push_back(fn.getPeriodic() ? 1 : 0);
        franges.push_back(lo);
        franges.push_back(hi);
        fvalues.insert(fvalues.end(), v.begin(), v.end());
    }
    const auto cp = cstr(pnames), cg = cstr(gnames), ci = cstr(inames), cvn = cstr(vnames), cve = cstr(vexprs),
               ct = cstr(texprs), cd = cstr(derivs_), cf = cstr(fnames);
    gbi_custom_spec spec{};
    spec.n_functions = (int32_t)fnames.size();
    spec.function_names = cf.data();
    spec.function_kinds = fkinds.data();
    spec.function_sizes = fsizes.data();
    spec.function_periodic = fperiodic.data();
    spec.function_ranges = franges.data();
    spec.function_values = fvalues.data();
    spec.function_dims = fdims.data();
    spec.function_ranges_yz = franges_yz.data();
    spec.n_params = np;
    spec.param_names = cp.data();
    spec.param_values = pvals.data();
    spec.radius_param = radius;
    spec.n_globals = (int32_t)gnames.size();
    spec.global_names = cg.data();
    spec.global_values = gvals.data();
    spec.n_integrals = (int32_t)inames.size();
    spec.integral_names = ci.data();
    spec.n_values = (int32_t)vnames.size();
    spec.value_names = cvn.data();
    spec.value_exprs = cve.data();
    spec.n_terms = (int32_t)texprs.size();
    spec.term_exprs = ct.data();
    spec.term_types = ttypes.data();
    spec.n_param_derivs = (int32_t)derivs_.size();
    spec.param_deriv_names = cd.data();
    gbi_destroy(ctx_);
    ctx_ = nullptr;
    n_ = n;
    periodic_ = p.nonbonded_method == GBI_CUTOFF_PERIODIC;
    check(gbi_create_custom(n, &p, &spec, &ctx_));
    if (excl && force.getNumExclusions() > 0) {
        std::vector<int32_t> pairs(2 * (size_t)force.getNumExclusions());
        for (int k = 0; k < force.getNumExclusions(); ++k) {
            int a, b;
            force.getExclusionParticles(k, a, b);
            pairs[2 * k] = a;
            pairs[2 * k + 1] = b;
        }
        check(gbi_set_exclusions(ctx_, 0, force.getNumExclusions(), pairs.data()));
    }
    custom_ = true;
}

void HipCalcCharmmGBMVForceKernel::initialize(const System& system, const CharmmGBMVForce& force) {
    if (system.getNumParticles() != force.getNumParticles())  // CharmmGBMVForceImpl.cpp:58-59
        throw OpenMMException("CharmmGBMVForce must have exactly as many particles as the System it belongs to.");
    globals_.clear();
    global_values_.clear();
    for (int k = 0; k < force.getNumGlobalParameters(); ++k) {
        globals_.push_back(force.getGlobalParameterName(k));
        global_values_.push_back(force.getGlobalParameterDefaultValue(k));
    }
    derivs_.clear();
    for (int k = 0; k < force.getNumEnergyParameterDerivatives(); ++k)
        derivs_.push_back(force.getEnergyParameterDerivativeName(k));
    if (force.getGBIntegralType() != CharmmGBMVForce::CustomIntegral && !isCanonicalGBMV(force)) {
        createCustom(force);  // general expressions: kernels generated from the force's expressions
        return;
    }
    custom_ = false;
    gbi_params p;
    std::vector<double> q, r;
    gbmvParamsFromForce(force, p, q, r);
    params_ = p;
    create(force.getNumParticles(), q, r, p);
    // a ParticlePair term skips the force's exclusions (CharmmReferenceGBMV.cpp:272-284)
    std::string expr;
    CharmmGBMVForce::ComputationType type;
    force.getEnergyTermParameters(1, expr, type);
    if (type == CharmmGBMVForce::ParticlePair && force.getNumExclusions() > 0) {
        std::vector<int32_t> pairs(2 * (size_t)force.getNumExclusions());
        for (int k = 0; k < force.getNumExclusions(); ++k) {
            int a, b;
            force.getExclusionParticles(k, a, b);
            pairs[2 * k] = a;
            pairs[2 * k + 1] = b;
        }
        check(gbi_set_exclusions(ctx_, 0, force.getNumExclusions(), pairs.data()));
    }
}

// The reference reads every global parameter from the Context at each execute
// (CharmmReferenceKernels.cpp:355-357); the engine gets the values that changed.
void HipCalcCharmmGBMVForceKernel::pullGlobals(ContextImpl& context) {
    bool changed = false;
    for (size_t k = 0; k < globals_.size(); ++k) {
        const double v = context.getParameter(globals_[k]);
        if (v != global_values_[k]) {
            global_values_[k] = v;
            changed = true;
        }
    }
    if (!changed) return;
    if (custom_) {
        check(gbi_set_custom_globals(ctx_, (int32_t)globals_.size(), global_values_.data()));
        return;
    }
    for (size_t k = 0; k < globals_.size(); ++k) {
        const std::string& g = globals_[k];
        const double v = global_values_[k];
        if (g == "alpha0") params_.alpha0 = v;
        else if (g == "alpha1") params_.alpha1 = v;
        else if (g == "eta") params_.eta = v;
        else if (g == "slope") params_.slope = v;
        else if (g == "shift") params_.shift = v;
        else if (g == "P") params_.prefactor = v;
    }
    check(gbi_update_constants(ctx_, 0, &params_));
}

double HipCalcCharmmGBMVForceKernel::execute(ContextImpl& context, bool includeForces, bool includeEnergy) {
    pullGlobals(context);
    const double e = run(context, includeForces, includeEnergy);
    if (custom_ && !derivs_.empty()) {  // CharmmReferenceKernels.cpp:360-368
        std::vector<double> d(derivs_.size());
        check(gbi_get_custom_param_derivs(ctx_, (int32_t)d.size(), d.data()));
        addEnergyParameterDerivatives(context, derivs_, d);
    }
    return e;
}

void HipCalcCharmmGBMVForceKernel::copyParametersToContext(ContextImpl& context, const CharmmGBMVForce& force) {
    if (custom_) {
        if (force.getNumParticles() != n_)  // CharmmReferenceKernels.cpp:373-374
            throw OpenMMException("updateParametersInContext: The number of particles has changed");
        // new per-particle parameters only (CharmmReferenceKernels.cpp:376-385): the generated
        // kernels stay
        const std::vector<double> pvals = particleParameterArray(force);
        check(gbi_update_custom_parameters(ctx_, n_, pvals.data()));
        context.systemChanged();
        return;
    }
    gbi_params p;
    std::vector<double> q, r;
    gbmvParamsFromForce(force, p, q, r);
    update(q, r);
    context.systemChanged();
}

}  // namespace gbi_openmm
