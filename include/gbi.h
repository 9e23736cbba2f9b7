/*
 * gbi.h -- C ABI of the MI355X (gfx950) integral Generalized-Born engine.
 *
 * This is the drop-in boundary for the hot path of ymwangg/OpenMMCustomGBIntegralForce:
 * the kernel interfaces CalcCharmmGBSWForceKernel and CalcCharmmGBMVForceKernel
 * (openmmapi/include/openmm/charmmKernels.h:49-129) reduced to plain pointers and sizes.
 * No exception crosses this boundary; every call returns a gbi_status and
 * gbi_last_error() describes the last failure of the calling thread.
 *
 * Units: nm, kJ/mol, proton charge.  Positions and forces are N x 3 row-major
 * doubles (the Reference platform's vector<Vec3> layout).  On the device, forces are
 * accumulated in 64-bit fixed point with scale 2^32 (the OpenMM CUDA force-buffer
 * convention), so results are bitwise reproducible.
 *
 * Threading: a context is driven by one host thread at a time; all device work of a
 * context is ordered on its HIP stream.
 *
 * Batches: a context holds S >= 1 independent systems (gbi_batch_create; gbi_create is
 * the S = 1 case).  Per-atom arrays of a batch are the systems' arrays concatenated in
 * system order (N = sum of the systems' atom counts); energies are one double per system.
 * Every pass of an evaluation is one launch over all systems.
 */
#ifndef GBI_H
#define GBI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBI_ABI_VERSION 6

typedef enum gbi_status {
    GBI_OK = 0,
    GBI_ERR_INVALID_ARGUMENT = 1,
    GBI_ERR_UNSUPPORTED = 2,       /* e.g. CustomIntegral (type 3), a triclinic box */
    GBI_ERR_PARTICLE_COUNT = 3,    /* updateParametersInContext with another N */
    GBI_ERR_HIP = 4,               /* a HIP runtime call failed */
    GBI_ERR_OUT_OF_MEMORY = 5,
    GBI_ERR_GRID_TOO_LARGE = 6,    /* lookup grid longer than 1000 nm on an axis */
    GBI_ERR_NONFINITE = 7,         /* NaN/Inf in the inputs or the result */
    GBI_ERR_NO_DEVICE = 8,
    GBI_ERR_CAPACITY = 9           /* internal: grid buffers grew; the call was redone */
} gbi_status;

/* Which reference force the context evaluates. */
typedef enum gbi_model {
    GBI_MODEL_GBSW = 0, /* CharmmGBSWForce (CharmmReferenceGBSW.cpp) */
    GBI_MODEL_GBMV = 1  /* CharmmGBMVForce with the canonical GB energy (CharmmReferenceGBMV.cpp) */
} gbi_model;

/* CharmmGBMVForce::GBIntegralType (CharmmGBMVForce.h:99-104). */
typedef enum gbi_integral_type {
    GBI_INTEGRAL_GBSW = 0,
    GBI_INTEGRAL_GBMV1 = 1,
    GBI_INTEGRAL_GBMV2 = 2,
    GBI_INTEGRAL_CUSTOM = 3 /* unimplemented in the reference; rejected */
} gbi_integral_type;

/* Which atoms a quadrature point sees. */
typedef enum gbi_lookup_policy {
    GBI_LOOKUP_REFERENCE = 0, /* the reference lookup grid of the model (default) */
    GBI_LOOKUP_BRUTE = 1      /* every atom (the reference's computeVolume paths); small N only */
} gbi_lookup_policy;

/*
 * CalcCharmmGBMVForceKernel::NonbondedMethod (charmmKernels.h:52-56).  With a cutoff, the
 * STILL pair terms count only pairs inside it: r <= c for GBI_MODEL_GBSW
 * (CharmmReferenceGBSW.cpp:246-247), r < c for GBI_MODEL_GBMV (CharmmReferenceGBMV.cpp:300-301);
 * the self terms always count.  GBI_MODEL_GBSW also truncates its chain rule: dB_i/dx_j
 * reaches only atoms j within the cutoff of i (CharmmReferenceGBSW.cpp:285-296), so its
 * forces are not the exact gradient; GBI_MODEL_GBMV's chain rule is complete.  Periodic
 * systems take minimum-image distances everywhere and tile their rectangular box with the
 * lookup grid (gbi_set_periodic_box).  The adjoint (GBI_ADJOINT) differentiates the energy
 * under every method (GBI_MODEL_GBSW's truncation shapes its forces, not dE/dR or dE/dtheta).
 */
typedef enum gbi_nonbonded_method {
    GBI_NO_CUTOFF = 0,
    GBI_CUTOFF_NON_PERIODIC = 1,
    GBI_CUTOFF_PERIODIC = 2
} gbi_nonbonded_method;

/*
 * All constants of the path.  gbi_params_default() fills the reference values; every
 * field may be changed before gbi_create().
 */
typedef struct gbi_params {
    int32_t model;          /* gbi_model */
    int32_t integral_type;  /* gbi_integral_type (GBMV model; GBSW model uses GBSW) */
    int32_t lookup_policy;  /* gbi_lookup_policy */
    int32_t nonbonded_method;
    double cutoff;          /* nm (ignored for GBI_NO_CUTOFF) */
    /* quadrature: n_radial Gauss-Legendre points on [r0, r1] x n_lebedev directions */
    int32_t n_radial;
    int32_t n_lebedev;      /* 6, 14, 26, 38 or 50 */
    double r0, r1;
    /* lookup grid */
    double grid_length;     /* nm */
    double buffer;          /* nm */
    int32_t cell_cap;       /* atoms kept per cell (<= 0: uncapped) */
    int32_t orders[2];      /* integral orders; the canonical Born map uses {4, 7} */
    /* GBSW smoothing: R' = rscale (R + roffset), switching half-width sw */
    double sw, rscale, roffset;
    /* GBMV-I: gamma0, beta, lambda; GBMV-II: beta, lambda, P1, P2, S0 */
    double gamma0, beta, lambda, P1, P2, S0;
    double vguard;          /* GBMV-II: |v|^2 below this means S = +inf (V = 1 for beta < 0, 0 for beta > 0) */
    /* Born map B = slope / (alpha0 (1/eta - I4) + alpha1 (1/(4 eta^4) - I7)^(1/4)) + shift */
    double alpha0, alpha1, eta, slope, shift;
    /* energy E = sum_i P q_i^2/(2 B_i) + sum_{i<j} P q_i q_j / f_ij */
    double prefactor;
    /* performance knobs (0 = automatic) */
    double tight_grid_length; /* GBSW model: cell edge of the exact neighbour grid, nm (0: 0.20) */
    int32_t block_volume;     /* threads per workgroup of the volume pass */
    int32_t block_force;      /* threads per workgroup of the chain-rule pass */
} gbi_params;

typedef struct gbi_stats {
    double n_points;      /* quadrature points with a non-empty neighbour list */
    double n_vol_pairs;   /* point/atom pairs scanned by the volume pass */
    double n_cells;       /* lookup cells */
    double n_entries;     /* cell list entries */
    double n_atoms;
    double n_quad;        /* quadrature points per atom */
    double n_active;      /* points with a chain-rule term (0 < V < 1 for GBSW) */
    double n_frc_pairs;   /* point/atom pairs the chain-rule pass visits (active points x their lists) */
    double slot_bytes;    /* bytes a binned point occupies in its cell's slots, averaged over the atoms:
                             4 (packed point; systems binning by bricks) or 16 (FP32 record and packed point) */
} gbi_stats;

/* Device view of a context's buffers (for callers that keep data on the GPU). */
typedef struct gbi_device_view {
    double* positions;    /* N x 3 doubles */
    int64_t* forces;      /* 3 x N fixed point (x[N] | y[N] | z[N]), scale 2^32 */
    double* energy;       /* S doubles */
    double* born_radii;   /* N doubles */
    double* integrals;    /* 2 x N doubles (integral-major) */
    void* stream;         /* hipStream_t of the context */
} gbi_device_view;

typedef struct gbi_ctx gbi_ctx;

/* Flags of gbi_compute. */
#define GBI_FORCES 1u
#define GBI_ENERGY 2u
#define GBI_ADJOINT 4u /* also dE/dR_j, dE/dq_j and dE/dtheta (gbi_get_adjoint) */

/*
 * Constants theta of a system whose energy derivatives the adjoint returns
 * (gbi_get_adjoint, GBI_N_THETA doubles per system, zero where the model has no such
 * constant).  No reference counterpart: the reference hard-codes these constants.
 */
typedef enum gbi_theta {
    GBI_THETA_PREFACTOR = 0, /* P of the STILL energy */
    GBI_THETA_ALPHA0,        /* Born map */
    GBI_THETA_ALPHA1,
    GBI_THETA_ETA,           /* Born map only (the quadrature range r0 stays fixed) */
    GBI_THETA_SLOPE,
    GBI_THETA_SHIFT,
    GBI_THETA_SW,            /* GBSW switching half-width w */
    GBI_THETA_RSCALE,        /* GBSW R' = rscale (R + roffset) */
    GBI_THETA_ROFFSET,
    GBI_THETA_GAMMA0,        /* GBMV-I */
    GBI_THETA_BETA,          /* GBMV-I / GBMV-II */
    GBI_THETA_LAMBDA,        /* GBMV-I / GBMV-II */
    GBI_THETA_P1,            /* GBMV-II */
    GBI_THETA_P2,
    GBI_THETA_S0,
    GBI_N_THETA
} gbi_theta;

/* Launch-shape knobs (gbi_autotune / gbi_set_tuning). */
typedef struct gbi_tuning {
    int32_t blocks_per_cu;    /* workgroups per compute unit of the per-cell passes */
    int32_t pair_columns;     /* columns per pair tile, in 64-atom blocks of pair_columns/64 (0: automatic) */
    double tight_grid_length; /* nm, cell of the exact GBSW grid (GBSW model systems) */
    double ms_per_eval;       /* gbi_autotune: measured time of the chosen shape */
} gbi_tuning;

/* Reference defaults of a model (constructors of GBSWIntegral / GBMVIntegralType{I,II},
 * CharmmReferenceGBSW.cpp:20-52). */
gbi_status gbi_params_default(int32_t model, int32_t integral_type, gbi_params* out);

/*
 * Create a context: replaces CalcCharmmGB{SW,MV}ForceKernel::initialize(System, Force)
 * (charmmKernels.h:71,111).  charge / radius: N doubles (copied).
 */
gbi_status gbi_create(int32_t n_atoms, const double* charge, const double* radius,
                      const gbi_params* params, gbi_ctx** out);

/*
 * A batch of n_systems independent systems in one context (no reference counterpart: the
 * reference evaluates one System per Context).  n_atoms[s] atoms each; charge and radius
 * hold sum(n_atoms) doubles in system order; params[s] configures system s (models,
 * integral types and quadratures may differ).
 */
gbi_status gbi_batch_create(int32_t n_systems, const int32_t* n_atoms, const double* charge, const double* radius,
                            const gbi_params* params, gbi_ctx** out);
gbi_status gbi_get_layout(gbi_ctx* ctx, int32_t* n_systems, int32_t* n_atoms_total);

/*
 * One evaluation: replaces CalcCharmmGB{SW,MV}ForceKernel::execute(context, includeForces,
 * includeEnergy) (charmmKernels.h:81,121).  pos: N x 3 host doubles.  Forces are ADDED to
 * `forces` (N x 3 host doubles, may be NULL when include_forces is 0); the energy of each
 * system is written to energy[0..S-1] (may be NULL).  Synchronous.
 */
gbi_status gbi_execute(gbi_ctx* ctx, const double* pos, int32_t include_forces, int32_t include_energy,
                       double* energy, double* forces);

/*
 * Replaces copyParametersToContext (charmmKernels.h:88,128): new charges and radii for the
 * same N; GBI_ERR_PARTICLE_COUNT when n_atoms differs.
 */
gbi_status gbi_update_parameters(gbi_ctx* ctx, int32_t n_atoms, const double* charge, const double* radius);

/*
 * New constants of system `system`: the global parameters a Context holds for the canonical
 * CharmmGBMVForce (alpha0, alpha1, eta, slope, shift, prefactor P, read by the reference from
 * the Context at every execute, CharmmReferenceKernels.cpp:355-357) and the volume constants
 * (sw, rscale, roffset, gamma0, beta, lambda, P1, P2, S0, vguard) a fitting loop moves.  Every
 * other field of `params` must equal the system's (GBI_ERR_INVALID_ARGUMENT otherwise).
 */
gbi_status gbi_update_constants(gbi_ctx* ctx, int32_t system, const gbi_params* params);

/*
 * The periodic box of a GBI_CUTOFF_PERIODIC system: box_vectors = 3 x 3 doubles (a, b, c
 * row-major), read by the reference kernels from the Context every evaluation
 * (GBSWIntegral.cpp:108-110, CharmmReferenceKernels.cpp:351, :430).  Rectangular only
 * (GBI_ERR_UNSUPPORTED otherwise, as CustomGBIntegral.cpp:47-50); every edge at least twice
 * the cutoff (CharmmGBMVForceImpl.cpp:99-104) and above twice the lookup grid's reach.
 * Must be called before the first evaluation of such a system and whenever the box changes.
 */
gbi_status gbi_set_periodic_box(gbi_ctx* ctx, int32_t system, const double* box_vectors);

/*
 * Excluded pairs of a GBI_MODEL_GBMV system whose pair energy term is ParticlePair
 * (CharmmGBMVForce::addExclusion; CharmmReferenceGBMV.cpp:272-284 skips them; a
 * ParticlePairNoExclusions term takes none): n_pairs pairs of local atom indices (2 x int32
 * each), replacing the system's previous list (n_pairs = 0 clears it).  The self terms and
 * the Born radii are unaffected.  GBI_MODEL_GBSW systems have no exclusions.
 */
gbi_status gbi_set_exclusions(gbi_ctx* ctx, int32_t system, int32_t n_pairs, const int32_t* pairs);

/*
 * General CharmmGBMVForce expressions (CharmmGBMVForce.h: addPerParticleParameter,
 * addGlobalParameter, addGBIntegral, addComputedValue, addEnergyTerm), Lepton syntax:
 * + - * / ^, unary minus, parentheses, sqrt exp log sin cos tan asin acos atan sinh cosh tanh
 * erf erfc abs step delta square cube recip min max select.  Computed values are
 * SingleParticle (as the reference evaluates them, CharmmReferenceGBMV.cpp:173-176) and may
 * use parameters, globals, integrals and earlier values; SingleParticle energy terms see the
 * same; pair terms (ParticlePair, ParticlePairNoExclusions) see r, globals and parameters and
 * values with suffixes 1 and 2 (atom1 < atom2).  Integrals: one or two, of orders
 * params->orders[0..n_integrals-1] (>= 2, GBSWIntegral.cpp:92-95), of params->integral_type.
 */
typedef struct gbi_custom_spec {
    int32_t n_params;
    const char* const* param_names;
    const double* param_values;      /* n_atoms x n_params, row-major */
    int32_t radius_param;            /* the parameter that is the atomic radius of the integrals */
    int32_t n_globals;
    const char* const* global_names;
    const double* global_values;
    int32_t n_integrals;             /* 1 or 2 */
    const char* const* integral_names;
    int32_t n_values;                /* 0..4 */
    const char* const* value_names;
    const char* const* value_exprs;
    int32_t n_terms;
    const char* const* term_exprs;
    const int32_t* term_types;       /* 0 SingleParticle, 1 ParticlePair, 2 ParticlePairNoExclusions */
    /* CharmmGBMVForce::addEnergyParameterDerivative (CharmmGBMVForce.h:258): up to 4 global
     * parameters whose energy derivative each evaluation also returns (gbi_get_custom_param_derivs) */
    int32_t n_param_derivs;
    const char* const* param_deriv_names;
    /* CharmmGBMVForce::addTabulatedFunction (CharmmGBMVForce.h:442; the reference evaluates them
     * with OpenMM's tabulated-function classes, CharmmReferenceKernels.cpp:187-191): functions
     * callable by name in every expression.  kinds: 0 Continuous1DFunction (a cubic spline
     * through sizes[k] values equally spaced on [ranges[2k], ranges[2k + 1]]; natural and zero
     * outside the range, or periodic[k] with period max - min and equal first and last values);
     * 1, 2, 3 Discrete1D/2D/3DFunction of 1, 2, 3 arguments (values[x + xsize (y + ysize z)]
     * at the rounded arguments, dims[3k..3k+2] = xsize, ysize, zsize; a non-finite result when
     * an index leaves its size); 4, 5 Continuous2D/3DFunction (values[x + xsize (y + ysize z)]
     * at equally spaced points of [ranges[2k], ranges[2k + 1]] x [ranges_yz[4k], ranges_yz[4k + 1]]
     * (x [ranges_yz[4k + 2], ranges_yz[4k + 3]]), at least 2 per axis; the bicubic (tricubic)
     * Hermite patch whose knot derivatives come from cubic splines along each axis, natural and
     * zero outside the ranges, or periodic on every axis).  values: every function's values,
     * concatenated. */
    int32_t n_functions;
    const char* const* function_names;
    const int32_t* function_kinds;
    const int32_t* function_sizes;
    const int32_t* function_periodic;
    const double* function_ranges;
    const double* function_values;
    const int32_t* function_dims;
    const double* function_ranges_yz;  /* 4 per function; read for the Continuous2D/3D kinds only */
} gbi_custom_spec;

/*
 * A one-system context for general expressions: replaces CalcCharmmGBMVForceKernel::initialize
 * (charmmKernels.h:71) for a force whose expressions are not the canonical GB forms.  The
 * expressions are compiled for the device here (parse errors: GBI_ERR_INVALID_ARGUMENT with
 * the message).  Evaluations, cutoffs, periodic boxes and exclusions work as for other
 * contexts; the adjoint and gbi_update_parameters are canonical-only.
 */
gbi_status gbi_create_custom(int32_t n_atoms, const gbi_params* params, const gbi_custom_spec* spec, gbi_ctx** out);
/* New values of the global parameters (Context::setParameter), n_globals doubles. */
gbi_status gbi_set_custom_globals(gbi_ctx* ctx, int32_t n_globals, const double* values);
/*
 * dE/d(global parameter) of the last evaluation, one double per spec->param_deriv_names entry:
 * the values the reference adds into the Context's energy-parameter derivatives after each
 * execute (CharmmReferenceKernels.cpp:360-368, CharmmReferenceGBMV.cpp:222-228, 263-264,
 * 345-346, 440-443).  The total derivative, through every computed value.
 */
gbi_status gbi_get_custom_param_derivs(gbi_ctx* ctx, int32_t n_derivs, double* out);
/*
 * New per-particle parameters of a general-expression context (n_atoms x n_params, row-major;
 * the same N): replaces CalcCharmmGBMVForceKernel::copyParametersToContext
 * (CharmmReferenceKernels.cpp:372-386) without recompiling the kernels.
 */
gbi_status gbi_update_custom_parameters(gbi_ctx* ctx, int32_t n_atoms, const double* param_values);
/* The computed values of the last evaluation: n_values x N doubles (value-major). */
gbi_status gbi_get_custom_values(gbi_ctx* ctx, double* values);

/* Asynchronous pieces of gbi_execute for device-resident callers. */
gbi_status gbi_set_positions(gbi_ctx* ctx, const double* pos, int32_t pos_on_device);
gbi_status gbi_compute(gbi_ctx* ctx, uint32_t flags);   /* enqueue one evaluation */
gbi_status gbi_synchronize(gbi_ctx* ctx);                /* wait + check device flags */
gbi_status gbi_get_energy(gbi_ctx* ctx, double* energy);
gbi_status gbi_get_forces(gbi_ctx* ctx, double* forces, int32_t add);
gbi_status gbi_get_born_radii(gbi_ctx* ctx, double* born);
gbi_status gbi_get_integrals(gbi_ctx* ctx, double* integrals); /* 2 x N (integral-major) */
gbi_status gbi_get_stats(gbi_ctx* ctx, gbi_stats* stats);
gbi_status gbi_get_device_view(gbi_ctx* ctx, gbi_device_view* view);

/*
 * The buffers of an OpenMM GPU platform, on the device and asynchronous on the context's
 * stream, in the platform's atom order: atom_slot (device, N int32, NULL for the identity)
 * gives the slot of each atom.  gbi_set_positions_posq reads positions from posq, one
 * (x, y, z, q) per slot as float4 (posq_double = 0) or double4 (1); gbi_add_forces_fixed
 * ADDS the last evaluation's forces, int64 fixed point with scale 2^32, to force[d *
 * padded_n + slot] for axis d (the layout into which the CUDA platform's integral kernel
 * adds its forces, reduceIntegralForce.cu:61-77, and its N^2 kernel, customGBEnergyN2.cu).
 */
gbi_status gbi_set_positions_posq(gbi_ctx* ctx, const void* posq, int32_t posq_double, const int32_t* atom_slot);
gbi_status gbi_add_forces_fixed(gbi_ctx* ctx, int64_t* force, int32_t padded_n, const int32_t* atom_slot);
gbi_status gbi_set_stream(gbi_ctx* ctx, void* hip_stream); /* NULL: the context's own stream */

/*
 * Adjoint of the last GBI_ADJOINT evaluation (no reference counterpart): dE/dR (N),
 * dE/dq (N) and dE/dtheta (S x GBI_N_THETA, gbi_theta order).  Any pointer may be NULL.
 * The lookup grids' membership tests are held fixed (the derivative of the evaluated
 * piece).  gbi_execute_adjoint = set positions + GBI_FORCES|GBI_ENERGY|GBI_ADJOINT + reads.
 */
gbi_status gbi_get_adjoint(gbi_ctx* ctx, double* dEdR, double* dEdq, double* dEdtheta);
gbi_status gbi_execute_adjoint(gbi_ctx* ctx, const double* pos, double* energy, double* forces, double* dEdR,
                               double* dEdq, double* dEdtheta);

/*
 * One evaluation of one NoCutoff system that allocates nothing on the device: gbi_create +
 * gbi_execute + gbi_destroy in one synchronous call, with host inputs (charge, radius: N
 * doubles; pos: N x 3) and outputs (energy; forces ADDED, N x 3).  Every device buffer of the
 * evaluation is carved from a static array of 256 MiB per GPU that the library holds from its
 * load until the process ends; the call makes no hipMalloc, hipFree or other allocating call.
 * Systems whose buffers need more (about 2,500 atoms at 24 x 38, 1,400 at 24 x 50) fail with
 * GBI_ERR_INVALID_ARGUMENT naming the size; other nonbonded methods with GBI_ERR_UNSUPPORTED.
 * The array keeps the last call's contents; calls in one process are serialised (a call that
 * overlapped another on the same GPU would race on the array).  hip_stream: NULL for the null
 * stream.  No counterpart in the reference: for callers that cannot allocate device memory
 * (serving loops and harnesses that forbid allocation).
 */
gbi_status gbi_execute_noalloc(int32_t n_atoms, const double* charge, const double* radius, const gbi_params* params,
                               const double* pos, int32_t include_forces, int32_t include_energy, double* energy,
                               double* forces, void* hip_stream);

/*
 * Replay evaluations from a captured hipGraph (one graph launch per evaluation instead of
 * ~25 kernel launches).  The graph is re-captured when buffers grow.  Not used while
 * profiling is enabled.
 */
gbi_status gbi_set_graph(gbi_ctx* ctx, int32_t enable);

/*
 * Launch shapes: gbi_autotune times `evals` evaluations of each candidate shape on the
 * current positions, keeps the fastest and reports it; gbi_set_tuning sets one directly.
 * Results do not depend on the shape except through the GBSW exact-grid cell length
 * (FP32 summation order of V inside a cell).
 */
gbi_status gbi_autotune(gbi_ctx* ctx, int32_t evals, gbi_tuning* out);
gbi_status gbi_set_tuning(gbi_ctx* ctx, const gbi_tuning* tuning);
gbi_status gbi_get_tuning(gbi_ctx* ctx, gbi_tuning* out);

/*
 * Diagnostics (no reference counterpart): HIP events recorded on the context's stream
 * around each pass of every evaluation while enabled.  Passes: 0 grid build, 1 point
 * binning, 2 volume, 3 integrals + Born radii, 4 pair energy, 5 dE/dB -> dE/dI + energy,
 * 6 chain-rule forces, 7 owner-force reduction.
 */
#define GBI_PROFILE_PASSES 8
gbi_status gbi_profile_enable(gbi_ctx* ctx, int32_t enable); /* enabling clears the record */
gbi_status gbi_profile_read(gbi_ctx* ctx, double* avg_ms, int32_t* launches); /* avg_ms[GBI_PROFILE_PASSES] */
const char* gbi_profile_pass_name(int32_t pass);

void gbi_destroy(gbi_ctx* ctx);
const char* gbi_last_error(void);
int32_t gbi_abi_version(void);

/*
 * Quadrature tables of the library (for inspection and tests): Q x 5 (x, y, z, r, w),
 * index = radial * n_lebedev + angular, as GBSWIntegral.cpp:41-56.
 */
gbi_status gbi_quadrature(int32_t n_radial, double r0, double r1, int32_t n_lebedev, double* out);

#ifdef __cplusplus
}
#endif
#endif /* GBI_H */
