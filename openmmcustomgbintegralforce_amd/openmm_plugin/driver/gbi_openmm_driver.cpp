// Runs one CharmmGBSWForce or canonical CharmmGBMVForce through the HIP kernels the way
// OpenMM does: the kernel is created by name from the "HIP" platform's registered factory,
// then initialize / execute / copyParametersToContext.  Used by the tests.
//
// Input (binary, native endianness): int32 model (0 CharmmGBSWForce, 1 CharmmGBMVForce with
// the canonical expressions, 2 CharmmGBMVForce with the general ones of generalGBMV, 3 the
// general ones of derivGBMV, which asks for energy parameter derivatives, 4 those of tabGBMV,
// with tabulated functions and the atoms' coordinates),
// int32 integral type, int32 Lebedev rule (0 = the reference's), int32 N, then N x 3
// positions, N charges, N radii (doubles), optionally followed by int32 nonbonded method,
// double cutoff and 3 doubles of a rectangular periodic box, then optionally int32 n and n
// excluded pairs (2 x int32) that make the GBMV pair term ParticlePair.  Output (text): energy,
// forces, the energy after copyParametersToContext with charges x 1.5, and the outcome of
// the error paths; model 1 adds the energy after Context::setParameter("alpha0", -0.2),
// model 3 the energy parameter derivatives, and the energy and dE/dk after
// Context::setParameter("k", 0.45).
//
// `gbi_openmm_driver --canonical-check` (no device needed) prints which CharmmGBMVForce
// variants the kernel class takes as the canonical forms.
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>

#include "HipCharmmKernels.h"
#include "openmm/Platform.h"

extern "C" void registerKernelFactories();

using namespace OpenMM;

namespace {

template <class T>
void rd(FILE* f, T* p, size_t n) {
    if (fread(p, sizeof(T), n, f) != n) throw std::runtime_error("short input");
}

CharmmGBMVForce* canonicalGBMV(const std::vector<double>& q, const std::vector<double>& r, int type, int rule,
                               const std::vector<int32_t>& excl) {
    auto* f = new CharmmGBMVForce();
    f->addPerParticleParameter("q");
    f->addPerParticleParameter("radius");
    f->addGlobalParameter("alpha0", -0.180);
    f->addGlobalParameter("alpha1", 1.817);
    f->addGlobalParameter("eta", 0.05);
    f->addGlobalParameter("slope", 1.0);
    f->addGlobalParameter("shift", 0.0);
    f->addGlobalParameter("P", -138.935456 * (1.0 - 1.0 / 80.0));
    std::vector<int> i4 = {4}, i7 = {7};
    if (rule > 0) {
        i4.push_back(rule);
        i7.push_back(rule);
    }
    f->addGBIntegral("I4", i4);
    f->addGBIntegral("I7", i7);
    f->setGBIntegralType((CharmmGBMVForce::GBIntegralType)type);
    f->addComputedValue("B", gbi_openmm::kCanonicalBornValue, CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm(gbi_openmm::kCanonicalSelfEnergy, CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm(gbi_openmm::kCanonicalPairEnergy,
                     excl.empty() ? CharmmGBMVForce::ParticlePairNoExclusions : CharmmGBMVForce::ParticlePair);
    for (size_t i = 0; i < q.size(); ++i) f->addParticle({q[i], r[i]});
    for (size_t k = 0; k + 1 < excl.size(); k += 2) f->addExclusion(excl[k], excl[k + 1]);
    return f;
}

// A non-canonical CharmmGBMVForce: one integral, a Born value of another form, the STILL pair
// term (the same expressions as tests/test_gpu_openmm_plugin.py's oracle call).
CharmmGBMVForce* generalGBMV(const std::vector<double>& q, const std::vector<double>& r, int type) {
    auto* f = new CharmmGBMVForce();
    f->addPerParticleParameter("q");
    f->addPerParticleParameter("radius");
    f->addGlobalParameter("P", -138.935456 * (1.0 - 1.0 / 80.0));
    f->addGBIntegral("I4", std::vector<int>{4});
    f->setGBIntegralType((CharmmGBMVForce::GBIntegralType)type);
    f->addComputedValue("B", "1/(1/radius-0.5*I4*radius)", CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm("0.5*P*q^2/B", CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm(gbi_openmm::kCanonicalPairEnergy, CharmmGBMVForce::ParticlePairNoExclusions);
    for (size_t i = 0; i < q.size(); ++i) f->addParticle({q[i], r[i]});
    return f;
}

// generalGBMV with a global k inside the Born value and energy derivatives with respect to k
// and P (CharmmGBMVForce::addEnergyParameterDerivative).
CharmmGBMVForce* derivGBMV(const std::vector<double>& q, const std::vector<double>& r, int type) {
    auto* f = new CharmmGBMVForce();
    f->addPerParticleParameter("q");
    f->addPerParticleParameter("radius");
    f->addGlobalParameter("P", -138.935456 * (1.0 - 1.0 / 80.0));
    f->addGlobalParameter("k", 0.5);
    f->addGBIntegral("I4", std::vector<int>{4});
    f->setGBIntegralType((CharmmGBMVForce::GBIntegralType)type);
    f->addComputedValue("B", "1/(1/radius-k*I4*radius)", CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm("0.5*P*q^2/B", CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm(gbi_openmm::kCanonicalPairEnergy, CharmmGBMVForce::ParticlePairNoExclusions);
    f->addEnergyParameterDerivative("k");
    f->addEnergyParameterDerivative("P");
    for (size_t i = 0; i < q.size(); ++i) f->addParticle({q[i], r[i]});
    return f;
}

// generalGBMV with Continuous1D/2D/3DFunctions and Discrete1D/2D/3DFunctions, and the coordinates x, y, z
// in the Born value and in the single-particle term (the same expressions as
// tests/test_gpu_openmm_plugin.py's oracle call).
CharmmGBMVForce* tabGBMV(const std::vector<double>& q, const std::vector<double>& r, int type) {
    auto* f = new CharmmGBMVForce();
    f->addPerParticleParameter("q");
    f->addPerParticleParameter("radius");
    f->addGlobalParameter("P", -138.935456 * (1.0 - 1.0 / 80.0));
    f->addGBIntegral("I4", std::vector<int>{4});
    f->setGBIntegralType((CharmmGBMVForce::GBIntegralType)type);
    std::vector<double> sv(31);
    for (int k = 0; k < 31; ++k) sv[k] = 1.0 + 0.2 * std::sin(0.3 * k);
    f->addTabulatedFunction("scale", new Continuous1DFunction(sv, 0.0, 3.0));
    f->addTabulatedFunction("w", new Discrete1DFunction({0.9, 1.1}));
    f->addTabulatedFunction("lut", new Discrete2DFunction(2, 3, {0.7, 0.8, 0.95, 1.05, 1.1, 1.2}));
    std::vector<double> cv(12);
    for (int k = 0; k < 12; ++k) cv[k] = 1.0 + 0.03 * k;
    f->addTabulatedFunction("tri3", new Discrete3DFunction(2, 2, 3, cv));
    std::vector<double> s2(5 * 4), v3(4 * 3 * 3);
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 5; ++i) s2[i + 5 * j] = 0.2 * std::sin(0.7 * i) + 0.1 * j;
    for (int k = 0; k < 3; ++k)
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 4; ++i) v3[i + 4 * (j + 3 * k)] = 1.0 + 0.05 * i - 0.03 * j + 0.02 * k * i;
    f->addTabulatedFunction("surf", new Continuous2DFunction(5, 4, s2, -1.5, 1.5, -1.5, 1.5));
    f->addTabulatedFunction("vol3", new Continuous3DFunction(4, 3, 3, v3, -1.5, 1.5, -1.5, 1.5, -1.5, 1.5));
    f->addComputedValue("B", "scale(10*radius)*lut(step(q), 10*radius)*(1+0.002*(x+y-z))/(1/radius-0.5*I4*radius)",
                        CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm("0.5*P*q^2/B*w(step(q))*tri3(step(q), step(radius-0.18), 20*radius-2)+0.01*q^2*(x^2+y^2+z^2)"
                     "+0.2*q^2*surf(x, y)+0.3*q^2*vol3(x, z, y)",
                     CharmmGBMVForce::SingleParticle);
    f->addEnergyTerm(gbi_openmm::kCanonicalPairEnergy, CharmmGBMVForce::ParticlePairNoExclusions);
    for (size_t i = 0; i < q.size(); ++i) f->addParticle({q[i], r[i]});
    return f;
}

// The Context holds every global parameter of its forces, starting at their defaults.
void setDefaults(ContextImpl& context, const CharmmGBMVForce& force) {
    for (int k = 0; k < force.getNumGlobalParameters(); ++k)
        context.setParameter(force.getGlobalParameterName(k), force.getGlobalParameterDefaultValue(k));
}

// The same evaluation with the Context's state on the device, as a GPU platform keeps it:
// posq (double4) and an int64 fixed-point force buffer, padded, atoms in reversed slots.
// Prints whether energy and forces equal the host path's bitwise.
template <class Exec>
void deviceRun(FILE* out, ContextImpl& context, const std::vector<Vec3>& pos, const std::vector<double>& q,
               const std::vector<Vec3>& host_forces, double host_energy, Exec exec) {
    const int n = (int)pos.size(), padded = (n + 31) / 32 * 32;
    std::vector<int> slot(n);
    std::vector<double> posq(4 * (size_t)padded, 0.0);
    for (int a = 0; a < n; ++a) {
        slot[a] = n - 1 - a;
        for (int d = 0; d < 3; ++d) posq[4 * (size_t)slot[a] + d] = pos[a][d];
        posq[4 * (size_t)slot[a] + 3] = q[a];
    }
    void *d_posq = nullptr, *d_force = nullptr, *d_slot = nullptr;
    auto ok = [](hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP: ") + hipGetErrorString(e));
    };
    ok(hipMalloc(&d_posq, sizeof(double) * posq.size()));
    ok(hipMalloc(&d_force, sizeof(long long) * 3 * padded));
    ok(hipMalloc(&d_slot, sizeof(int) * n));
    ok(hipMemcpy(d_posq, posq.data(), sizeof(double) * posq.size(), hipMemcpyHostToDevice));
    ok(hipMemset(d_force, 0, sizeof(long long) * 3 * padded));
    ok(hipMemcpy(d_slot, slot.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    DeviceBuffers& db = context.deviceBuffers();
    db.posq = d_posq;
    db.posq_double = 1;
    db.force = static_cast<long long*>(d_force);
    db.padded_n = padded;
    db.atom_slot = static_cast<const int*>(d_slot);
    const double e = exec();
    ok(hipDeviceSynchronize());
    std::vector<long long> f(3 * (size_t)padded);
    ok(hipMemcpy(f.data(), d_force, sizeof(long long) * f.size(), hipMemcpyDeviceToHost));
    int equal = 1;
    for (int a = 0; a < n; ++a)
        for (int d = 0; d < 3; ++d)
            equal &= (double)f[(size_t)d * padded + slot[a]] / 4294967296.0 == host_forces[a][d];
    fprintf(out, "device_energy_equal %d\ndevice_forces_equal %d\n", (int)(e == host_energy), equal);
    db = DeviceBuffers();
    ok(hipFree(d_posq));
    ok(hipFree(d_force));
    ok(hipFree(d_slot));
}

// Which forces the kernel class takes as canonical (host only).
int canonicalCheck() {
    const std::vector<double> q = {0.5, -0.5}, r = {0.2, 0.17};
    std::unique_ptr<CharmmGBMVForce> canon(canonicalGBMV(q, r, 2, 0, {}));
    // the integral named I4 of order 7 and the one named I7 of order 4: the Born expression
    // names them, so this force is not the canonical one
    std::unique_ptr<CharmmGBMVForce> swapped(canonicalGBMV(q, r, 2, 0, {}));
    swapped->setGBIntegralParameters(0, "I4", std::vector<int>{7}, std::vector<double>());
    swapped->setGBIntegralParameters(1, "I7", std::vector<int>{4}, std::vector<double>());
    // the same integrals declared in the other order are still canonical
    std::unique_ptr<CharmmGBMVForce> reordered(canonicalGBMV(q, r, 2, 0, {}));
    reordered->setGBIntegralParameters(0, "I7", std::vector<int>{7}, std::vector<double>());
    reordered->setGBIntegralParameters(1, "I4", std::vector<int>{4}, std::vector<double>());
    std::unique_ptr<CharmmGBMVForce> deriv(canonicalGBMV(q, r, 2, 0, {}));
    deriv->addEnergyParameterDerivative("alpha0");
    printf("canonical %d\nswapped_names %d\nreordered %d\nwith_param_derivative %d\n",
           (int)gbi_openmm::isCanonicalGBMV(*canon), (int)gbi_openmm::isCanonicalGBMV(*swapped),
           (int)gbi_openmm::isCanonicalGBMV(*reordered), (int)gbi_openmm::isCanonicalGBMV(*deriv));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::string(argv[1]) == "--canonical-check") return canonicalCheck();
    if (argc != 3) {
        fprintf(stderr, "usage: %s input.bin output.txt\n", argv[0]);
        return 2;
    }
    try {
        FILE* in = fopen(argv[1], "rb");
        if (!in) throw std::runtime_error("cannot open input");
        int32_t hdr[4];
        rd(in, hdr, 4);
        const int model = hdr[0], type = hdr[1], rule = hdr[2], n = hdr[3];
        std::vector<double> x(3 * (size_t)n), q(n), r(n);
        rd(in, x.data(), x.size());
        rd(in, q.data(), q.size());
        rd(in, r.data(), r.size());
        int32_t method = 0;
        double cutoff = 1.0, box[3] = {2.0, 2.0, 2.0};
        std::vector<int32_t> excl;
        if (fread(&method, sizeof(method), 1, in) == 1) {
            rd(in, &cutoff, 1);
            rd(in, box, 3);
            int32_t nex = 0;
            if (fread(&nex, sizeof(nex), 1, in) == 1 && nex > 0) {
                excl.resize(2 * (size_t)nex);
                rd(in, excl.data(), excl.size());
            }
        }
        fclose(in);

        Platform::registerPlatform(new Platform("HIP"));
        registerKernelFactories();
        Platform& platform = Platform::getPlatformByName("HIP");
        System system;
        for (int i = 0; i < n; ++i) system.addParticle(12.0);
        ContextImpl context(n);
        std::vector<Vec3> pos(n);
        for (int i = 0; i < n; ++i) pos[i] = Vec3(x[3 * i], x[3 * i + 1], x[3 * i + 2]);
        context.setPositions(pos);
        context.setPeriodicBoxVectors(Vec3(box[0], 0.0, 0.0), Vec3(0.0, box[1], 0.0), Vec3(0.0, 0.0, box[2]));

        FILE* out = fopen(argv[2], "w");
        if (!out) throw std::runtime_error("cannot open output");
        double e = 0.0, e_scaled = 0.0, e_noforce = 0.0;
        std::vector<double> qs(q);
        for (double& v : qs) v *= 1.5;
        if (model == 0) {
            CharmmGBSWForce force;
            force.setNonbondedMethod((CharmmGBSWForce::NonbondedMethod)method);
            force.setCutoffDistance(cutoff);
            for (int i = 0; i < n; ++i) force.addParticle(q[i], r[i], 1.0);
            std::unique_ptr<KernelImpl> k(platform.createKernel(CalcCharmmGBSWForceKernel::Name(), context));
            auto& kernel = dynamic_cast<CalcCharmmGBSWForceKernel&>(*k);
            kernel.initialize(system, force);
            e = kernel.execute(context, true, true);
            std::vector<Vec3> f = context.getForces();
            fprintf(out, "energy %.17g\n", e);
            for (int i = 0; i < n; ++i) fprintf(out, "%.17g %.17g %.17g\n", f[i][0], f[i][1], f[i][2]);
            deviceRun(out, context, pos, q, f, e, [&] { return kernel.execute(context, true, true); });
            e_noforce = kernel.execute(context, false, true);
            for (int i = 0; i < n; ++i) force.setParticleParameters(i, qs[i], r[i], 1.0);
            kernel.copyParametersToContext(context, force);
            e_scaled = kernel.execute(context, false, true);
            CharmmGBSWForce shorter;
            for (int i = 0; i + 1 < n; ++i) shorter.addParticle(q[i], r[i], 1.0);
            int count_throws = 0;
            try {
                kernel.copyParametersToContext(context, shorter);
            } catch (const OpenMMException&) {
                count_throws = 1;
            }
            fprintf(out, "count_change_throws %d\n", count_throws);
        } else {
            auto make = [&](const std::vector<double>& charges) {
                if (model == 4) return tabGBMV(charges, r, type);
                return model == 3 ? derivGBMV(charges, r, type)
                                  : (model == 2 ? generalGBMV(charges, r, type) : canonicalGBMV(charges, r, type, rule, excl));
            };
            std::unique_ptr<CharmmGBMVForce> force(make(q));
            force->setNonbondedMethod((CharmmGBMVForce::NonbondedMethod)method);
            force->setCutoffDistance(cutoff);
            setDefaults(context, *force);
            std::unique_ptr<KernelImpl> k(platform.createKernel(CalcCharmmGBMVForceKernel::Name(), context));
            auto& kernel = dynamic_cast<CalcCharmmGBMVForceKernel&>(*k);
            kernel.initialize(system, *force);
            e = kernel.execute(context, true, true);
            std::vector<Vec3> f = context.getForces();
            fprintf(out, "energy %.17g\n", e);
            for (int i = 0; i < n; ++i) fprintf(out, "%.17g %.17g %.17g\n", f[i][0], f[i][1], f[i][2]);
            if (model == 1) deviceRun(out, context, pos, q, f, e, [&] { return kernel.execute(context, true, true); });
            if (model == 3) {
                // one execute adds dE/dk and dE/dP once; the second (with k = 0.45) adds again
                auto& d = context.energyParameterDerivatives();
                fprintf(out, "dEdk %.17g\ndEdP %.17g\n", d["k"], d["P"]);
                context.setParameter("k", 0.45);
                d.clear();
                const double e2 = kernel.execute(context, false, true);
                fprintf(out, "energy_k045 %.17g\ndEdk_k045 %.17g\n", e2, d["k"]);
                context.setParameter("k", 0.5);
            }
            if (model == 1) {
                // a global parameter changed in the Context reaches the next execute
                context.setParameter("alpha0", -0.2);
                fprintf(out, "energy_alpha0 %.17g\n", kernel.execute(context, false, true));
                context.setParameter("alpha0", -0.180);
            }
            e_noforce = kernel.execute(context, false, true);
            std::unique_ptr<CharmmGBMVForce> scaled(make(qs));
            scaled->setNonbondedMethod((CharmmGBMVForce::NonbondedMethod)method);
            scaled->setCutoffDistance(cutoff);
            kernel.copyParametersToContext(context, *scaled);
            e_scaled = kernel.execute(context, false, true);
            fprintf(out, "count_change_throws 1\n");
        }
        fprintf(out, "energy_no_forces %.17g\n", e_noforce);
        fprintf(out, "energy_scaled %.17g\n", e_scaled);
        int illegal = 0;
        try {
            platform.createKernel("CalcSomethingElse", context);
        } catch (const OpenMMException&) {
            illegal = 1;
        }
        fprintf(out, "illegal_name_throws %d\n", illegal);
        fclose(out);
    } catch (const std::exception& ex) {
        fprintf(stderr, "gbi_openmm_driver: %s\n", ex.what());
        return 1;
    }
    return 0;
}
