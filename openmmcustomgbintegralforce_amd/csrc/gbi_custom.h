// General CharmmGBMVForce expressions: computed values, single-particle and pair energy
// terms given as expression strings (the Lepton syntax of CharmmGBMVForce::addComputedValue /
// addEnergyTerm).  The expressions are translated to HIP source in which every quantity is a
// forward-mode dual number, so one evaluation yields the value and the derivatives the chain
// rule needs; the source is compiled with hiprtc for the device when the context is created.
// Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace gbi {

constexpr int kMaxCustomValues = 4;
constexpr int kMaxCustomDerivs = 4;  // energy derivatives with respect to global parameters
constexpr int kMaxCustomIntegrals = 2;
constexpr int kCustomPairWaves = 4;  // waves of a cg_pairs workgroup (one 64 x 64 tile)
constexpr int kCustomChainWaves = 4; // waves of a cg_chain workgroup (one 64-atom block)

enum CustomTermType : int { TERM_SINGLE = 0, TERM_PAIR = 1, TERM_PAIR_NO_EXCL = 2 };

// Tabulated functions (CharmmGBMVForce::addTabulatedFunction with OpenMM's
// Continuous1D/2D/3DFunction and Discrete1D/2D/3DFunction).  Continuous: values at equally
// spaced points of [min, max] per axis (values[x + xsize (y + ysize z)]); one argument, the
// cubic spline through them; two or three, the bicubic (tricubic) Hermite patch whose knot
// derivatives d/dx, d/dy, d2/dxdy, ... come from cubic splines along each axis.  Natural
// splines and zero outside the ranges, or periodic with period max - min per axis.
// Discrete: the value at the rounded arguments, each rounded index inside its size.
enum TabKind : int {
    TAB_CONTINUOUS_1D = 0,
    TAB_DISCRETE_1D = 1,
    TAB_DISCRETE_2D = 2,
    TAB_DISCRETE_3D = 3,
    TAB_CONTINUOUS_2D = 4,
    TAB_CONTINUOUS_3D = 5
};
struct TabFunction {
    std::string name;
    int kind = TAB_CONTINUOUS_1D;
    bool periodic = false;
    double min = 0.0, max = 1.0;                          // x range of the continuous kinds
    double ymin = 0.0, ymax = 1.0, zmin = 0.0, zmax = 1.0;  // Continuous2D/3D: y and z ranges
    int dims[3] = {0, 1, 1};  // tables of 2 or 3 arguments: xsize, ysize, zsize
    std::vector<double> values;
};
inline int tab_arity(int kind) {
    if (kind == TAB_DISCRETE_2D || kind == TAB_CONTINUOUS_2D) return 2;
    if (kind == TAB_DISCRETE_3D || kind == TAB_CONTINUOUS_3D) return 3;
    return 1;
}
// Second derivatives of the natural (periodic) cubic spline through `values` at equally spaced
// points h apart; periodic splines need values.front() == values.back().  Throws on bad input.
std::vector<double> spline_second_derivatives(const std::vector<double>& values, double h, bool periodic);
// First derivatives of that spline at its knots.
std::vector<double> spline_knot_derivatives(const std::vector<double>& values, double h, bool periodic);
// The knot tables of a Continuous2D/3DFunction: {f, f_x, f_y, f_xy} or
// {f, f_x, f_y, f_z, f_xy, f_xz, f_yz, f_xyz}, each indexed like the values.  Throws on bad input.
std::vector<std::vector<double>> hermite_tables(const TabFunction& f);

struct CustomSpec {
    std::vector<std::string> params;     // per-particle parameter names
    std::vector<std::string> globals;    // global parameter names
    std::vector<std::string> integrals;  // integral names (1 or 2)
    std::vector<std::string> value_names, value_exprs;
    std::vector<std::string> term_exprs;
    std::vector<int> term_types;
    std::vector<int> deriv_globals;      // globals whose energy derivative is computed (indices)
    std::vector<TabFunction> functions;  // tabulated functions, callable by name in every expression
};

// Compiled kernels of one spec.  Kernel signatures (all arrays device, double unless noted;
// ND = the number of parameter derivatives, arrays sized with max(ND, 1)):
//   cg_values(int N, prm[N x NP], integ[2 x N], glob[NG], vals[NV x N], dvdi[NV x 2 x N],
//             esp[N], despdi[2 x N], dvdth[NV x ND x N], despdth[ND x N], pos[N x 3],
//             dvdx[NV x 3 x N], despdx[3 x N])                        grid N/256, block 256
//   cg_pairs(int N, int nb, pos[N x 3], prm, vals, glob, int method, double cut2,
//            const double* box3, const int* ex_ptr, const int* ex_idx,
//            part[nb (nb + 1) / 2 tiles x NPL x 64])     grid (nb, nb), block 64 kCustomPairWaves
//   cg_chain(int N, int nb, part, dvdi, esp, despdi, atom_energy[N], long long* force_fix[3 x N],
//            dedi[2 x N], dvdth, despdth, pdth[ND x N], dvdx, despdx)
//                                                          grid nb, block 64 kCustomChainWaves
// (nb = ceil(N / 64) atom blocks; NPL = n_planes planes per tile: 4 + NV + ND row, 3 + NV
// column; pdth holds each atom's share of dE/dtheta, summed over atoms on the host; dvdx and
// despdx are the derivatives of the values and single-particle terms with respect to the
// atom's own coordinates x, y, z, which computed values and single-particle terms may use)
struct CustomProgram {
    hipModule_t module = nullptr;
    hipFunction_t values = nullptr, pairs = nullptr, chain = nullptr;
    int n_params = 0, n_globals = 0, n_values = 0, n_planes = 0, n_derivs = 0;
    bool any_pair = false;
    std::string source;
    CustomProgram() = default;
    CustomProgram(const CustomProgram&) = delete;
    CustomProgram& operator=(const CustomProgram&) = delete;
    ~CustomProgram();
};

// Translates one expression to HIP source over the dual type D.  context: 0 computed value
// number `index` (sees parameters, globals, integrals, earlier values and x, y, z), 1
// single-particle energy term (also every value), 2 pair energy term (sees r, parameters and
// values with suffixes 1 and 2, globals); tabulated functions everywhere.  Throws
// std::runtime_error naming the problem.
std::string translate_expression(const std::string& expr, const CustomSpec& spec, int context, int index);

// Validates the spec and generates the kernels' source (host only, no device); throws
// std::runtime_error with the parser's message.
void generate_custom_source(const CustomSpec& spec, CustomProgram& out);
// Compiles out.source for `arch` (e.g. "gfx950") and loads the kernels; throws
// std::runtime_error with the compiler's message.
void compile_custom_program(const std::string& arch, CustomProgram& out);

}  // namespace gbi
