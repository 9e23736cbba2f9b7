// Host side of the C ABI (include/gbi.h): parameter defaults, the batch layout (per-system
// ranges of atoms, quadrature rows, cells and point slots), device buffers, capacity
// management, graph capture and the execute sequence.  Replaces the kernel drivers
// ReferenceCalcCharmmGB{SW,MV}ForceKernel::execute (CharmmReferenceKernels.cpp:337-370,
// :426-437) and CudaCalcCharmmGBMVForceKernel::execute / CustomGBIntegral::evaluate
// (CharmmCudaKernels.cpp:845-993, cuda CustomGBIntegral.cpp:625-717).
#include "gbi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gbi_custom.h"
#include "gbi_internal.h"
#include "quadrature.h"

namespace {

thread_local std::string g_last_error;

gbi_status fail(gbi_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(e_ == hipErrorOutOfMemory ? GBI_ERR_OUT_OF_MEMORY : GBI_ERR_HIP,           \
                        std::string(#expr) + ": " + hipGetErrorString(e_));                        \
    } while (0)

#define ST_TRY(expr)                      \
    do {                                  \
        gbi_status s_ = (expr);           \
        if (s_ != GBI_OK) return s_;      \
    } while (0)

constexpr double kOne4PiEps0 = 138.935456;  // OpenMM 7.x SimTKOpenMMRealType.h
constexpr double kTightMargin = 1e-6;       // nm, guards the exact GBSW lists against rounding
constexpr long long kMaxCells = 16ll * 1024 * 1024;  // limit of the three-pass scan
constexpr int kDefaultBlocksPerCU = 128;  // about one workgroup per occupied cell: the dispatcher balances them
// Cell edge of the exact GBSW grid (lists are exact at any edge): 0.20 nm measured fastest at
// 5k and 25k atoms (0.15 nm: 14 % slower, more distinct cells to bin; 0.30 nm: longer lists).
constexpr double kDefaultTightGrid = 0.20;
constexpr int kCutoffTileBlocks = 3;  // column blocks per pair tile of a cutoff system
constexpr double kSigmoidFp64Above = 16.0;  // |beta lambda| above which GBMV sums S in FP64
constexpr size_t kOwnAllocation = 2u << 20;  // bytes: a buffer not sub-allocated beside others

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    // min_bytes: a floor on the allocation's size (2 MB gives the buffer allocations of its own)
    hipError_t ensure(size_t count, size_t min_bytes = 0) {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        const size_t cnt = std::max<size_t>({count, 1, (min_bytes + sizeof(T) - 1) / sizeof(T)});
        hipError_t e = hipMalloc((void**)&p, cnt * sizeof(T));
        if (e == hipSuccess) n = cnt;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// Host description of one system of the batch.
struct SysHost {
    gbi_params P{};
    int N = 0, atom_begin = 0, Q = 0, quad_base = 0, pt_base = 0;
    gbi::GridKind kind = gbi::GRID_FAITHFUL;
    gbi::VolKind vol = gbi::VOL_GBSW;
    double g = 0.15, gpad = 0.0, halfdiag = 0.0, max_reach = 0.0;
    long long cell_cap = 0;
    int pt_cap = 1;
    int tmp_cap = 0;  // raw member slots per cell: estimated from the positions, grown from the device counts
    int pair_splits = 1, pair_k = 1, nblk = 1, blk_base = 0;
    long long colpart_base = 0;
    double box[3] = {0.0, 0.0, 0.0};  // periodic box edges (gbi_set_periodic_box)
    bool box_set = false;
    std::vector<int32_t> excl;         // excluded pairs, local atom indices (gbi_set_exclusions)
    bool custom = false;               // general expressions (gbi_create_custom)
    bool periodic() const { return P.nonbonded_method == GBI_CUTOFF_PERIODIC; }
    bool cutoff() const { return P.nonbonded_method != GBI_NO_CUTOFF; }
};

// Binning bricks: sk x sk x sk cells of about 0.45 nm reserve their point slots together (an
// atom's 1,200 GBMV-II points fall into about 200 such bricks against 670 cells of 0.15 nm;
// 0.17 ms for binning 5,000 atoms against 0.21 by cells).  Grids of cells above 0.18 nm bin by
// cells (sk = 1): on the 0.20 nm exact GBSW grid, bricks of 2 or 3 cells per edge were slower
// (0.49-0.52 ms against 0.44 for 25,000 atoms), their second pass costing more than the
// atomics they save.
constexpr double kBrickEdgeNm = 0.45;
int brick_edge(const SysHost& s) {
    if (s.kind == gbi::GRID_BRUTE) return 1;
    const int sk = (int)std::lround(kBrickEdgeNm / s.g);
    return sk >= 3 ? std::min(gbi::kMaxBrickEdge, sk) : 1;
}

// Cells per axis rounded up to whole bricks.
long long brick_cells(long long n, int sk) { return (n + sk - 1) / sk * sk; }

}  // namespace

struct gbi_ctx {
    std::vector<SysHost> sys;
    int S = 0, N = 0, NQ = 0, quad_rows = 0;
    std::vector<double> charge, radius;
    long long cell_total = 0;
    long long prec_total = 0;
    int entry_cap = 0;
    bool kinds[gbi::VOL_KINDS] = {false, false, false};
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    int num_cus = 256;
    bool have_positions = false;
    bool computed = false;
    bool computed_adjoint = false;
    gbi_tuning tuning{};

    // device buffers
    DevBuf<gbi::SysDesc> d_sys;
    DevBuf<gbi::SysGeom> geom;
    DevBuf<int> flags_all, atom_sys, cell_sys, bin_order, bbox_done;
    DevBuf<double> bbox_part;
    DevBuf<int4> pair_tiles;
    DevBuf<int> colblk;
    DevBuf<double> d_box;
    DevBuf<float4> blk_lo, blk_hi;
    DevBuf<int> pair_atom;
    std::vector<int> pair_order_host;  // host copy of pair_atom
    DevBuf<int> excl_start, excl_col;
    DevBuf<unsigned long long> excl_mask;
    bool any_excl = false;

    // general expressions (one system)
    std::unique_ptr<gbi::CustomProgram> cprog;
    int custom_nparams = 0, custom_radius = 0;
    gbi::CustomLaunch claunch;
    DevBuf<double> c_prm, c_glob, c_vals, c_dvdi, c_esp, c_despdi, c_part, c_dvdth, c_despdth, c_pdth, c_dvdx,
        c_despdx;
    DevBuf<int> c_ex_ptr, c_ex_idx;
    bool any_cut = false, any_trunc = false;
    DevBuf<float> colpart;
    int n_tiles = 0, max_splits = 1, n_colblocks = 0;
    long long colpart_total = 1;
    DevBuf<double> pos, d_charge, d_radius, d_reach, quad_off, quad_pw, integrals, born, dbdi, dedi, pair_part,
        atom_energy, energy, stats, dedq, dedr, theta_atom, theta;
    DevBuf<float4> aparam;
    DevBuf<int> cell_fill, cell_start, cell_count_eff, cell_atoms_tmp, cell_atoms, cell_pt_count,
        cell_act_count, frec_base, worklist, worklist_sys, work, scan_tmp;
    DevBuf<float4> ent_pos, ent_par, prec, frec0, frec1, frec2, vrec;
    DevBuf<int> brick_count, cell_ctr;
    DevBuf<uint2> brick_rec;
    int brick_total = 0, brick_max = 0;
    bool any_bricks = false, any_cells = true;
    DevBuf<float> quad_r2, vsig, vval;
    DevBuf<long long> force_fix, dedr_fix, vacc_fix;

    // graph replay (index: adjoint evaluation)
    bool use_graph = false;
    hipGraphExec_t graph[2] = {nullptr, nullptr};

    // diagnostics: kPasses + 1 events per profiled evaluation
    bool profiling = false;
    std::vector<std::array<hipEvent_t, gbi::kPasses + 1>> prof_events;
    size_t prof_used = 0;

    void drop_graphs() {
        for (auto& g : graph)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }

    void release_profile() {
        for (auto& a : prof_events)
            for (auto& e : a) (void)hipEventDestroy(e);
        prof_events.clear();
        prof_used = 0;
    }

    ~gbi_ctx() {
        drop_graphs();
        release_profile();
        d_sys.release();
        geom.release();
        DevBuf<double>* d[] = {&pos,       &d_charge, &d_radius, &d_reach, &quad_off,   &quad_pw, &integrals,
                               &born,      &dbdi,     &dedi,     &pair_part, &atom_energy, &energy, &stats,
                               &dedq,      &dedr,     &theta_atom, &theta};
        for (auto* b : d) b->release();
        aparam.release();
        pair_tiles.release();
        colblk.release();
        d_box.release();
        blk_lo.release();
        blk_hi.release();
        pair_atom.release();
        excl_start.release();
        excl_col.release();
        excl_mask.release();
        DevBuf<double>* cb[] = {&c_prm,     &c_glob,    &c_vals, &c_dvdi, &c_esp,   &c_despdi,
                                &c_part,    &c_dvdth,   &c_despdth, &c_pdth, &c_dvdx, &c_despdx};
        for (auto* b : cb) b->release();
        c_ex_ptr.release();
        c_ex_idx.release();
        cprog.reset();
        colpart.release();
        bbox_part.release();
        bbox_done.release();
        DevBuf<int>* ib[] = {&flags_all,      &atom_sys,     &cell_sys,       &bin_order,       &cell_fill,
                             &cell_start,     &cell_count_eff, &cell_atoms_tmp, &cell_atoms, &cell_pt_count,
                             &cell_act_count, &frec_base,    &worklist,       &worklist_sys, &work, &scan_tmp};
        for (auto* b : ib) b->release();
        DevBuf<float4>* fb[] = {&ent_pos, &ent_par, &prec, &frec0, &frec1, &frec2, &vrec};
        for (auto* b : fb) b->release();
        brick_count.release();
        cell_ctr.release();
        brick_rec.release();
        quad_r2.release();
        vsig.release();
        vval.release();
        force_fix.release();
        dedr_fix.release();
        vacc_fix.release();
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

gbi_status validate_params(const gbi_params& P, int N, bool custom = false) {
    if (N <= 0) return fail(GBI_ERR_INVALID_ARGUMENT, "n_atoms must be positive");
    if (P.model != GBI_MODEL_GBSW && P.model != GBI_MODEL_GBMV)
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown model");
    if (P.model == GBI_MODEL_GBMV && P.integral_type == GBI_INTEGRAL_CUSTOM)
        return fail(GBI_ERR_UNSUPPORTED, "CustomIntegral (type 3) is not implemented by the reference either");
    if (P.model == GBI_MODEL_GBMV && (P.integral_type < 0 || P.integral_type > 2))
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown GB integral type");
    if (P.nonbonded_method < GBI_NO_CUTOFF || P.nonbonded_method > GBI_CUTOFF_PERIODIC)
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown nonbonded method");
    if (P.nonbonded_method != GBI_NO_CUTOFF && !(P.cutoff > 0.0))
        return fail(GBI_ERR_INVALID_ARGUMENT, "the cutoff distance must be positive");
    if (P.nonbonded_method == GBI_CUTOFF_PERIODIC && P.lookup_policy == GBI_LOOKUP_BRUTE)
        return fail(GBI_ERR_UNSUPPORTED, "the brute-force lookup has no periodic form");
    if (P.n_radial <= 0 || P.n_radial * P.n_lebedev > (1 << gbi::kQBits) - 1)
        return fail(GBI_ERR_INVALID_ARGUMENT, "quadrature must have 1..2047 points per atom");
    if (!(P.n_lebedev == 6 || P.n_lebedev == 14 || P.n_lebedev == 26 || P.n_lebedev == 38 || P.n_lebedev == 50))
        return fail(GBI_ERR_INVALID_ARGUMENT, "n_lebedev must be 6, 14, 26, 38 or 50");
    if (!(P.r0 > 0.0) || !(P.r1 > P.r0)) return fail(GBI_ERR_INVALID_ARGUMENT, "need 0 < r0 < r1");
    if (!(P.grid_length > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "grid_length must be positive");
    if (P.model == GBI_MODEL_GBSW || P.integral_type == GBI_INTEGRAL_GBSW) {
        if (!(P.sw > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "the switching half-width must be positive");
    } else {
        // V = 1/(1 + exp(beta (S - lambda))): beta = 0 makes V independent of the geometry
        if (!std::isfinite(P.beta) || P.beta == 0.0 || !std::isfinite(P.lambda))
            return fail(GBI_ERR_INVALID_ARGUMENT, "GBMV beta must be finite and nonzero, lambda finite");
    }
    for (int k = 0; k < 2; ++k)
        if (P.orders[k] < 2)  // GBSWIntegral.cpp:92-95
            return fail(GBI_ERR_INVALID_ARGUMENT, "the order of an integral must be >= 2");
    const bool o47 = (P.orders[0] == 4 && P.orders[1] == 7) || (P.orders[0] == 7 && P.orders[1] == 4);
    if (!o47 && !custom) return fail(GBI_ERR_UNSUPPORTED, "the canonical Born map needs integrals of order 4 and 7");
    return GBI_OK;
}

gbi_status check_atoms(int n, const double* charge, const double* radius) {
    for (int j = 0; j < n; ++j) {
        if (!std::isfinite(charge[j]) || !std::isfinite(radius[j]))
            return fail(GBI_ERR_NONFINITE, "non-finite charge or radius");
        if (!(radius[j] > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "radii must be positive");
    }
    return GBI_OK;
}

// Per-atom constants (float for the volume functions, double reach for the exact grid).
void atom_params(const gbi_ctx& c, std::vector<float4>& ap, std::vector<double>& reach) {
    ap.resize(c.N);
    reach.resize(c.N);
    for (const SysHost& s : c.sys) {
        const gbi_params& P = s.P;
        for (int j = s.atom_begin; j < s.atom_begin + s.N; ++j) {
            const double R = c.radius[j];
            float4 v = make_float4(0.f, 0.f, 0.f, (float)R);
            if (s.vol == gbi::VOL_GBSW) {
                const double Rp = (R + P.roffset) * P.rscale;
                const double lo = Rp - P.sw, hi = Rp + P.sw;
                v.x = (float)Rp;
                v.y = lo > 0.0 ? (float)(lo * lo) : -1.0f;
                v.z = (float)(hi * hi);
                reach[j] = hi + kTightMargin;
            } else if (s.vol == gbi::VOL_GBMV1) {
                v.x = (float)(P.gamma0 * std::log(P.lambda) / (R * R * R * R));
                reach[j] = 0.0;
            } else {
                const double C = P.P1 * R * 10.0 + P.P2;
                v.x = (float)C;
                v.y = (float)(C - R * R * 100.0);
                v.z = (float)(400.0 * C * C);
                reach[j] = 0.0;
            }
            ap[j] = v;
        }
    }
}

// Largest reach of each system's atoms (the tight grid's padding).
void set_max_reach(gbi_ctx& c, const std::vector<double>& reach) {
    for (SysHost& s : c.sys) {
        s.max_reach = 0.0;
        for (int j = s.atom_begin; j < s.atom_begin + s.N; ++j) s.max_reach = std::max(s.max_reach, reach[j]);
    }
}

gbi_status upload_atoms(gbi_ctx& c) {
    std::vector<float4> ap;
    std::vector<double> reach;
    atom_params(c, ap, reach);
    HIP_TRY(hipMemcpyAsync(c.aparam.p, ap.data(), sizeof(float4) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_reach.p, reach.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_charge.p, c.charge.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_radius.p, c.radius.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    set_max_reach(c, reach);
    HIP_TRY(hipStreamSynchronize(c.stream));
    return GBI_OK;
}

// System constants in the device layout.
gbi::SysDesc make_desc(const gbi_ctx& c, const SysHost& s, long long cell_base, long long prec_base,
                       long long tmp_base) {
    gbi::SysDesc d{};
    const gbi_params& P = s.P;
    d.grid_kind = s.kind;
    d.vol_kind = s.vol;
    d.atom_begin = s.atom_begin;
    d.atom_end = s.atom_begin + s.N;
    d.Q = s.Q;
    d.quad_base = s.quad_base;
    d.pt_base = s.pt_base;
    d.cell_base = (int)cell_base;
    d.cell_cap = (int)s.cell_cap;
    d.pt_cap = s.pt_cap;
    d.tmp_base = tmp_base;
    d.tmp_cap = s.tmp_cap;
    d.prec_base = prec_base;
    d.sk = brick_edge(s);
    d.cap = (P.cell_cap > 0 && s.kind != gbi::GRID_BRUTE) ? P.cell_cap : INT_MAX;
    d.pair_splits = s.pair_splits;
    d.pair_k = s.pair_k;
    d.nblk = s.nblk;
    d.colpart_base = s.colpart_base;
    d.i4 = P.orders[0] == 4 ? 0 : 1;
    d.i7 = 1 - d.i4;
    d.periodic = s.periodic() ? 1 : 0;
    d.cut_mode = !s.cutoff() ? 0 : (P.model == GBI_MODEL_GBSW ? 1 : 2);
    d.chain_trunc = (s.cutoff() && P.model == GBI_MODEL_GBSW) ? 1 : 0;
    d.blk_base = s.blk_base;
    d.custom = s.custom ? 1 : 0;
    d.cut2 = s.cutoff() ? P.cutoff * P.cutoff : 0.0;
    d.cut2_lo = s.cutoff() ? (float)(d.cut2 * (1.0 - 1e-4)) : FLT_MAX;
    d.cut2_hi = s.cutoff() ? (float)(d.cut2 * (1.0 + 1e-4)) : FLT_MAX;
    // GBSW shell atoms lie within max(R' + w) of a point (max_reach); 1e-4 nm absorbs FP32 rounding
    d.trunc_in = d.chain_trunc ? (float)(std::sqrt(d.cut2 * (1.0 - 1e-4)) - s.max_reach - 1e-4) : FLT_MAX;
    d.trunc_out = (d.chain_trunc && !s.periodic()) ? (float)(std::sqrt(d.cut2 * (1.0 + 1e-4)) + s.max_reach + 1e-4)
                                                   : FLT_MAX;
    d.g = s.g;
    d.inv_g = 1.0 / s.g;
    d.gpad = s.kind == gbi::GRID_FAITHFUL ? s.gpad : s.max_reach;
    d.halfdiag = s.halfdiag;
    d.buffer = P.buffer;
    d.k1 = (float)(3.0 / (4.0 * P.sw));
    d.k3 = (float)(1.0 / (4.0 * P.sw * P.sw * P.sw));
    d.sw = (float)P.sw;
    d.inv_sw = P.sw > 0.0 ? (float)(1.0 / P.sw) : 0.f;
    d.rscale = (float)P.rscale;
    d.roffset = (float)P.roffset;
    d.gamma0 = (float)P.gamma0;
    d.beta = (float)P.beta;
    d.lambda = (float)P.lambda;
    d.S0 = (float)P.S0;
    d.vguard = (float)P.vguard;
    d.fclamp = 1.0e6f;
    d.P1 = (float)P.P1;
    d.v_empty = s.vol == gbi::VOL_GBSW    ? 1.0f
                : s.vol == gbi::VOL_GBMV2 ? 0.0f
                                          : (float)(1.0 / (1.0 + std::exp(P.beta * (0.0 - P.lambda))));
    d.sign_chain = s.vol == gbi::VOL_GBSW ? 1.0f : -1.0f;
    // V = 1/(1 + exp(beta (S - lambda))) moves by |beta| S eps_S for a relative error eps_S of
    // S, and its derivative by about as much relative to itself: the FP32 sums' eps_S ~ 3e-7
    // stays well inside the force tolerance up to |beta lambda| ~ 16 (the reference constants
    // have 10); steeper or later sigmoids sum S in FP64
    d.acc64 = (s.vol != gbi::VOL_GBSW && std::fabs(P.beta * P.lambda) > kSigmoidFp64Above) ? 1 : 0;
    d.beta64 = P.beta;
    d.lambda64 = P.lambda;
    d.S064 = P.S0;
    d.alpha0 = P.alpha0;
    d.alpha1 = P.alpha1;
    d.eta = P.eta;
    d.slope = P.slope;
    d.shift = P.shift;
    d.prefactor = P.prefactor;
    (void)c;
    return d;
}

// Cell, slot and worklist buffers for the systems' current capacities; uploads the
// system descriptors and the cell -> system map.
gbi_status layout(gbi_ctx& c) {
    long long cells = 0, slots = 0, raw = 0, bricks = 0, brick_max = 0, brick_slots = 0;
    bool any_bricks = false, any_cells = false;
    for (const SysHost& s : c.sys) {
        cells += s.cell_cap;
        slots += s.cell_cap * (long long)s.pt_cap;
        raw += s.cell_cap * (long long)s.tmp_cap;
        const int sk = brick_edge(s);
        any_cells |= sk == 1;
        if (sk == 1) continue;  // binned by cells
        bricks += s.cell_cap / (sk * sk * sk);
        brick_max = std::max(brick_max, s.cell_cap / (sk * sk * sk));
        any_bricks = true;
    }
    // brick slots share the prec slot layout (a system binning by cells leaves its range unused)
    brick_slots = any_bricks ? slots : 1;
    if (cells > kMaxCells) return fail(GBI_ERR_GRID_TOO_LARGE, "more than 16M lookup cells in the batch");
    if (slots * (long long)sizeof(float4) > (200ll << 30))
        return fail(GBI_ERR_GRID_TOO_LARGE, "point slots of the lookup grids exceed 200 GB");
    if (raw * (long long)sizeof(int) > (64ll << 30))
        return fail(GBI_ERR_GRID_TOO_LARGE, "raw cell lists of the lookup grids exceed 64 GB");
    c.drop_graphs();
    HIP_TRY(hipStreamSynchronize(c.stream));
    c.cell_total = cells;
    c.prec_total = slots;
    HIP_TRY(c.cell_fill.ensure(cells));
    HIP_TRY(c.cell_start.ensure(cells + 1));
    HIP_TRY(c.cell_count_eff.ensure(cells));
    // the binning pass's atomic counters get allocations of their own: sub-allocated beside
    // other small buffers, that pass was measured 15-20 % slower in some processes (0.26 vs
    // 0.215 ms at 5k atoms, 0.49 vs 0.445 ms at 25k), never with its own allocation
    HIP_TRY(c.cell_pt_count.ensure(cells, kOwnAllocation));
    // counters on a line per 128 cells for the cells way (k_point_bin<false>: 0.419 -> 0.412 ms at
    // 25,000 GBSW atoms, 9,284 -> 9,959 conformation-evals/s for 8 x 2,000); bricks-only contexts
    // count in cell_pt_count, their split gaining nothing from it
    if (any_cells) HIP_TRY(c.cell_ctr.ensure((cells + 4095) / 4096 * 4096, kOwnAllocation));
    HIP_TRY(c.brick_count.ensure(std::max<long long>(1, bricks) * gbi::kBrickStride, kOwnAllocation));
    HIP_TRY(c.brick_rec.ensure(brick_slots));
    c.brick_total = (int)bricks;
    c.brick_max = (int)brick_max;
    c.any_bricks = any_bricks;
    c.any_cells = any_cells;
    HIP_TRY(c.cell_act_count.ensure(cells));
    HIP_TRY(c.frec_base.ensure(cells + 1));
    HIP_TRY(c.worklist.ensure(gbi::VOL_KINDS * cells));
    HIP_TRY(c.worklist_sys.ensure(gbi::VOL_KINDS * cells));
    HIP_TRY(c.cell_sys.ensure(cells));
    HIP_TRY(c.scan_tmp.ensure((cells + 4095) / 4096 + 16));
    HIP_TRY(c.prec.ensure(slots));
    HIP_TRY(c.cell_atoms_tmp.ensure(std::max<long long>(1, raw)));
    std::vector<int> cs(cells);
    std::vector<gbi::SysDesc> desc(c.S);
    long long cb = 0, pb = 0, tb = 0, bb = 0;
    for (int s = 0; s < c.S; ++s) {
        desc[s] = make_desc(c, c.sys[s], cb, pb, tb);
        desc[s].brick_base = (int)bb;
        if (desc[s].sk > 1) bb += c.sys[s].cell_cap / (desc[s].sk * desc[s].sk * desc[s].sk);
        std::fill(cs.begin() + cb, cs.begin() + cb + c.sys[s].cell_cap, s);
        cb += c.sys[s].cell_cap;
        pb += c.sys[s].cell_cap * (long long)c.sys[s].pt_cap;
        tb += c.sys[s].cell_cap * (long long)c.sys[s].tmp_cap;
    }
    HIP_TRY(hipMemcpyAsync(c.cell_sys.p, cs.data(), sizeof(int) * cells, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_sys.p, desc.data(), sizeof(gbi::SysDesc) * c.S, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return GBI_OK;
}

// Pair tiles: row block I x column blocks [J0, J1) of the same system, J >= I (each
// unordered pair once).  K column blocks per tile: `pair_columns` / 64 when set, else the
// smallest K that keeps about 8192 tiles over the batch (every system's share of the block
// pairs).  Column partials: one entry per (row block, atom) of each system.
void tile_layout(gbi_ctx& c, std::vector<int4>& tiles, std::vector<int>& colblk) {
    long long pairs_total = 0;
    for (const SysHost& s : c.sys) {
        const long long nb = (s.N + 63) / 64;
        pairs_total += nb * (nb + 1) / 2;
    }
    const long long target = 8192;
    tiles.clear();
    colblk.clear();
    c.max_splits = 1;
    long long cp = 0;
    for (SysHost& s : c.sys) {
        const int nb = (s.N + 63) / 64;
        int K;
        if (c.tuning.pair_columns > 0) {
            K = std::max(1, c.tuning.pair_columns / 64);
        } else {
            const long long pairs = (long long)nb * (nb + 1) / 2;
            const double share = (double)target * pairs / (double)pairs_total;  // tiles for this system
            K = (int)std::ceil((double)pairs / std::max(1.0, share));
        }
        // cutoff systems: most block pairs are skipped and the work gathers near the diagonal,
        // so narrow tiles spread it (3 blocks measured best at 25k atoms, cutoff 0.5-1.6 nm)
        if (c.tuning.pair_columns <= 0 && s.cutoff()) K = std::min(K, kCutoffTileBlocks);
        K = std::max(1, std::min(K, nb));
        s.pair_k = K;
        s.nblk = nb;
        s.pair_splits = (nb + K - 1) / K;  // pieces of row block 0, the most of any block
        s.colpart_base = cp;
        s.blk_base = (int)colblk.size();
        cp += 64ll * nb * (nb - 1) / 2;  // row blocks I < J for every column block J
        c.max_splits = std::max(c.max_splits, s.pair_splits);
        for (int I = 0; I < nb; ++I) {
            colblk.push_back(s.atom_begin + 64 * I);
            for (int J0 = I, k = 0; J0 < nb; J0 += K, ++k) {
                const int J1 = std::min(nb, J0 + K);
                tiles.push_back(make_int4(s.atom_begin + 64 * I, s.atom_begin + 64 * J0,
                                          s.atom_begin + std::min(s.N, 64 * J1), k));
            }
        }
    }
    c.n_tiles = (int)tiles.size();
    c.n_colblocks = (int)colblk.size();
    c.colpart_total = std::max<long long>(1, cp);
}

gbi_status build_tiles(gbi_ctx& c) {
    std::vector<int4> tiles;
    std::vector<int> colblk;
    tile_layout(c, tiles, colblk);
    c.drop_graphs();
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(c.pair_tiles.ensure(tiles.size()));
    HIP_TRY(c.colblk.ensure(colblk.size()));
    HIP_TRY(c.blk_lo.ensure(colblk.size()));
    HIP_TRY(c.blk_hi.ensure(colblk.size()));
    HIP_TRY(c.pair_part.ensure((size_t)c.max_splits * 6 * c.N));
    HIP_TRY(c.colpart.ensure((size_t)5 * c.colpart_total));
    HIP_TRY(hipMemcpyAsync(c.pair_tiles.p, tiles.data(), sizeof(int4) * tiles.size(), hipMemcpyHostToDevice,
                           c.stream));
    HIP_TRY(hipMemcpyAsync(c.colblk.p, colblk.data(), sizeof(int) * colblk.size(), hipMemcpyHostToDevice,
                           c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return GBI_OK;
}

// Cell count and list entries of one system for the given positions, with headroom.
void estimate_caps(const gbi_ctx& c, const SysHost& s, const double* pos, long long& cells, long long& entries,
                   int& raw) {
    if (s.kind == gbi::GRID_BRUTE) {
        cells = 1;
        entries = s.N;
        raw = s.N;
        return;
    }
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = pos[d];
    for (int a = 0; a < s.N; ++a)
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], pos[3 * a + d]);
            mx[d] = std::max(mx[d], pos[3 * a + d]);
        }
    const double g = s.g;
    // headroom: 0.5 nm per side plus 10 %
    const double gridpad = s.kind == gbi::GRID_FAITHFUL ? s.gpad : s.max_reach;
    cells = 1;
    for (int d = 0; d < 3; ++d) {
        // periodic: the box's cells (ceil(L/g) + 1 per axis), with room for a 5 % larger box;
        // whole bricks of cells
        const double len = s.periodic() ? s.box[d] * 1.05 : (mx[d] - mn[d]) * 1.1 + 2.0 * gridpad + 1.0;
        cells *= brick_cells((long long)std::ceil(len / g) + 2, brick_edge(s));
    }
    // entries: box of candidate cells per atom (an upper bound of the members)
    double side;
    if (s.kind == gbi::GRID_FAITHFUL && s.periodic()) {
        // samples of the periodic walk per axis: 2 (pad + g') / step + 1, pad = R + sqrt3/2 g' + buffer,
        // g' the largest and step the smallest cell edge L / (ceil(L / g) + 1)
        double maxR = 0.0, gmax = 0.0, gmin = 1e300;
        for (int j = s.atom_begin; j < s.atom_begin + s.N; ++j) maxR = std::max(maxR, c.radius[j]);
        for (int d = 0; d < 3; ++d) {
            const double st = s.box[d] / (std::ceil(s.box[d] / g) + 1.0);
            gmax = std::max(gmax, st);
            gmin = std::min(gmin, st);
        }
        side = std::floor(2.0 * (maxR + std::sqrt(3.0) / 2.0 * gmax + s.P.buffer + gmax) / gmin) + 3.0;
    } else if (s.kind == gbi::GRID_FAITHFUL) {
        double maxR = 0.0;
        for (int j = s.atom_begin; j < s.atom_begin + s.N; ++j) maxR = std::max(maxR, c.radius[j]);
        side = std::floor(2.0 * (maxR + s.halfdiag + s.P.buffer) / g) + 3.0;
    } else {
        side = std::floor(2.0 * s.max_reach / g) + 3.0;
    }
    entries = (long long)(side * side * side * s.N * 1.05) + 1024;
    // raw members of a cell: twice the mean density over the bounding box (or the periodic
    // box) times the candidate box of a cell; an overflow is sized from the device counts
    double vol = 1.0;
    for (int d = 0; d < 3; ++d) vol *= s.periodic() ? s.box[d] : std::max(mx[d] - mn[d], g);
    const double cand = side * g;
    const double est = 2.0 * s.N / vol * cand * cand * cand + 32.0;
    raw = (int)std::min<double>(est, (double)s.N * (s.periodic() ? 8.0 : 1.0) + 32.0);
}

// Pair slots of a cutoff system along a serpentine path: columns of 0.45 nm x 0.45 nm in x
// and y, walked in z, alternating direction so the path stays connected (positions wrapped
// into a periodic box).  64 consecutive slots are then a short compact run, so whole block
// pairs beyond the cutoff are skipped (for a 25k-atom globule at 1.6 nm: 27 % of the block
// pairs evaluated, against 37 % for a Morton order on 0.25 nm cells and all for a random
// order).  Other systems keep the identity.  Any order gives the same pairs; the order only
// decides which blocks can be skipped.
void pair_order(const SysHost& s, const double* pos, int* out) {
    for (int k = 0; k < s.N; ++k) out[k] = s.atom_begin + k;
    if (!s.cutoff() || s.N <= 64) return;
    constexpr double kSlab = 0.45;
    std::vector<double> x(3 * (size_t)s.N);
    double mn[2] = {0.0, 0.0};
    for (int k = 0; k < s.N; ++k)
        for (int d = 0; d < 3; ++d) {
            double v = pos[3 * k + d];
            if (s.periodic() && s.box[d] > 0.0) v -= s.box[d] * std::floor(v / s.box[d]);
            x[3 * k + d] = v;
            if (d < 2) mn[d] = k == 0 ? v : std::min(mn[d], v);
        }
    struct Key {
        long long ix, iy;
        double z;
        int atom;
    };
    std::vector<Key> key(s.N);
    for (int k = 0; k < s.N; ++k) {
        const long long ix = (long long)std::floor((x[3 * k] - mn[0]) / kSlab);
        const long long iy = (long long)std::floor((x[3 * k + 1] - mn[1]) / kSlab);
        key[k] = {ix, ix % 2 == 0 ? iy : -iy, (ix + iy) % 2 == 0 ? x[3 * k + 2] : -x[3 * k + 2], s.atom_begin + k};
    }
    std::sort(key.begin(), key.end(), [](const Key& a, const Key& b) {
        if (a.ix != b.ix) return a.ix < b.ix;
        if (a.iy != b.iy) return a.iy < b.iy;
        if (a.z != b.z) return a.z < b.z;
        return a.atom < b.atom;
    });
    for (int k = 0; k < s.N; ++k) out[k] = key[k].atom;
}

// Exclusion masks in pair-slot space: one entry per (row block, column block >= row block)
// holding excluded pairs, 64 row masks each (diagonal blocks hold both orders of a pair).
gbi_status build_exclusions(gbi_ctx& c) {
    std::vector<int> slot(c.N);
    for (int k = 0; k < c.N; ++k) slot[c.pair_order_host[k]] = k;
    std::map<std::pair<int, int>, int> entry;
    std::vector<unsigned long long> masks;
    for (const SysHost& s : c.sys)
        for (size_t k = 0; k + 1 < s.excl.size(); k += 2) {
            int sa = slot[s.atom_begin + s.excl[k]] - s.atom_begin, sb = slot[s.atom_begin + s.excl[k + 1]] - s.atom_begin;
            if (sa / 64 > sb / 64) std::swap(sa, sb);
            const std::pair<int, int> key(s.blk_base + sa / 64, s.blk_base + sb / 64);
            auto it = entry.find(key);
            if (it == entry.end()) {
                it = entry.emplace(key, (int)(masks.size() / 64)).first;
                masks.resize(masks.size() + 64, 0ull);
            }
            unsigned long long* m = masks.data() + 64 * (size_t)it->second;
            m[sa % 64] |= 1ull << (sb % 64);
            if (key.first == key.second) m[sb % 64] |= 1ull << (sa % 64);
        }
    std::vector<int> start(c.n_colblocks + 1, 0), col;
    std::vector<unsigned long long> sorted;
    col.reserve(entry.size());
    sorted.reserve(masks.size());
    for (const auto& kv : entry) {  // std::map: ordered by (row block, column block)
        start[kv.first.first + 1]++;
        col.push_back(kv.first.second);
        sorted.insert(sorted.end(), masks.begin() + 64 * (size_t)kv.second, masks.begin() + 64 * (size_t)kv.second + 64);
    }
    for (int b = 0; b < c.n_colblocks; ++b) start[b + 1] += start[b];
    c.drop_graphs();
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(c.excl_start.ensure(start.size()));
    HIP_TRY(c.excl_col.ensure(std::max<size_t>(1, col.size())));
    HIP_TRY(c.excl_mask.ensure(std::max<size_t>(1, sorted.size())));
    HIP_TRY(hipMemcpyAsync(c.excl_start.p, start.data(), sizeof(int) * start.size(), hipMemcpyHostToDevice, c.stream));
    if (!col.empty()) {
        HIP_TRY(hipMemcpyAsync(c.excl_col.p, col.data(), sizeof(int) * col.size(), hipMemcpyHostToDevice, c.stream));
        HIP_TRY(hipMemcpyAsync(c.excl_mask.p, sorted.data(), sizeof(unsigned long long) * sorted.size(),
                               hipMemcpyHostToDevice, c.stream));
    }
    HIP_TRY(hipStreamSynchronize(c.stream));
    c.any_excl = !col.empty();
    return GBI_OK;
}

// Excluded pairs of a general-expression system as sorted partner lists (CSR, both directions).
gbi_status build_custom_exclusions(gbi_ctx& c) {
    const SysHost& s = c.sys[0];
    std::vector<std::vector<int>> partners(s.N);
    for (size_t k = 0; k + 1 < s.excl.size(); k += 2) {
        partners[s.excl[k]].push_back(s.excl[k + 1]);
        partners[s.excl[k + 1]].push_back(s.excl[k]);
    }
    std::vector<int> ptr(s.N + 1, 0), idx;
    for (int a = 0; a < s.N; ++a) {
        std::sort(partners[a].begin(), partners[a].end());
        partners[a].erase(std::unique(partners[a].begin(), partners[a].end()), partners[a].end());
        ptr[a + 1] = ptr[a] + (int)partners[a].size();
        idx.insert(idx.end(), partners[a].begin(), partners[a].end());
    }
    c.drop_graphs();
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(c.c_ex_ptr.ensure(ptr.size()));
    HIP_TRY(c.c_ex_idx.ensure(std::max<size_t>(1, idx.size())));
    HIP_TRY(hipMemcpyAsync(c.c_ex_ptr.p, ptr.data(), sizeof(int) * ptr.size(), hipMemcpyHostToDevice, c.stream));
    if (!idx.empty())
        HIP_TRY(hipMemcpyAsync(c.c_ex_idx.p, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    c.claunch.ex_ptr = c.c_ex_ptr.p;
    c.claunch.ex_idx = c.c_ex_idx.p;
    return GBI_OK;
}

// Grow the systems' capacities for the device positions (first call, new radii, overflow).
gbi_status size_from_device(gbi_ctx& c, long long min_entries) {
    std::vector<double> host(3 * (size_t)c.N);
    HIP_TRY(hipMemcpyAsync(host.data(), c.pos.p, sizeof(double) * 3 * c.N, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    for (double v : host)
        if (!std::isfinite(v)) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    bool changed = c.cell_total == 0;
    long long entries = 0;
    if (c.any_cut) {
        std::vector<int> order(c.N);
        for (const SysHost& s : c.sys)
            pair_order(s, host.data() + 3 * (size_t)s.atom_begin, order.data() + s.atom_begin);
        HIP_TRY(hipMemcpyAsync(c.pair_atom.p, order.data(), sizeof(int) * c.N, hipMemcpyHostToDevice, c.stream));
        HIP_TRY(hipStreamSynchronize(c.stream));
        c.pair_order_host = order;
        if (c.any_excl) ST_TRY(build_exclusions(c));
    }
    for (SysHost& s : c.sys) {
        long long cells, ent;
        int raw;
        estimate_caps(c, s, host.data() + 3 * (size_t)s.atom_begin, cells, ent, raw);
        entries += ent;
        if (raw > s.tmp_cap) {
            s.tmp_cap = raw;
            changed = true;
        }
        if (cells > s.cell_cap) {
            s.cell_cap = cells;
            changed = true;
        }
    }
    entries = std::max(entries, min_entries);
    if (entries > INT_MAX - 1) return fail(GBI_ERR_GRID_TOO_LARGE, "lookup grid too large for the device buffers");
    if (entries > c.entry_cap) {
        c.drop_graphs();
        HIP_TRY(hipStreamSynchronize(c.stream));
        c.entry_cap = (int)entries;
        HIP_TRY(c.cell_atoms.ensure(entries));
        HIP_TRY(c.ent_pos.ensure(entries));
        HIP_TRY(c.ent_par.ensure(entries));
    }
    if (changed) return layout(c);
    return GBI_OK;
}

gbi::KParams make_kparams(gbi_ctx& c, bool adjoint) {
    gbi::KParams k{};
    k.S = c.S;
    k.N = c.N;
    k.NQ = c.NQ;
    k.cell_total = (int)c.cell_total;
    k.entry_cap = c.entry_cap;
    k.n_tiles = c.n_tiles;
    k.max_splits = c.max_splits;
    k.n_colblocks = c.n_colblocks;
    k.colpart_total = c.colpart_total;
    k.adjoint = adjoint ? 1 : 0;
    k.bin_order = c.bin_order.p;
    k.persist_blocks = c.num_cus * std::max(1, c.tuning.blocks_per_cu);
    k.sys = c.d_sys.p;
    k.geom = c.geom.p;
    k.box = c.d_box.p;
    k.any_cut = c.any_cut ? 1 : 0;
    k.any_trunc = c.any_trunc ? 1 : 0;
    k.blk_lo = c.blk_lo.p;
    k.blk_hi = c.blk_hi.p;
    k.pair_atom = c.pair_atom.p;
    k.custom = c.cprog ? &c.claunch : nullptr;
    k.any_excl = c.any_excl ? 1 : 0;
    k.excl_start = c.excl_start.p;
    k.excl_col = c.excl_col.p;
    k.excl_mask = c.excl_mask.p;
    k.flags_all = c.flags_all.p;
    k.bbox_part = c.bbox_part.p;
    k.bbox_done = c.bbox_done.p;
    k.bbox_slices = 1;
    for (const SysHost& s : c.sys) k.bbox_slices = std::max(k.bbox_slices, gbi::bbox_slices(s.N));
    k.atom_sys = c.atom_sys.p;
    k.cell_sys = c.cell_sys.p;
    k.pos = c.pos.p;
    k.charge = c.d_charge.p;
    k.radius = c.d_radius.p;
    k.reach = c.d_reach.p;
    k.aparam = c.aparam.p;
    k.quad_off = c.quad_off.p;
    k.quad_pw = c.quad_pw.p;
    k.quad_r2 = c.quad_r2.p;
    k.quad_rows = c.quad_rows;
    k.pair_tiles = c.pair_tiles.p;
    k.colblk = c.colblk.p;
    k.cell_fill = c.cell_fill.p;
    k.cell_start = c.cell_start.p;
    k.cell_count_eff = c.cell_count_eff.p;
    k.cell_atoms_tmp = c.cell_atoms_tmp.p;
    k.cell_atoms = c.cell_atoms.p;
    k.ent_pos = c.ent_pos.p;
    k.ent_par = c.ent_par.p;
    k.cell_pt_count = c.cell_pt_count.p;
    // bricks-only contexts and the no-alloc path count in cell_pt_count directly
    const bool perm = c.any_cells && c.cell_ctr.p;
    k.cell_ctr = perm ? c.cell_ctr.p : c.cell_pt_count.p;
    k.ctr_perm = perm ? 1 : 0;
    k.ctr_len = (int)(perm ? (c.cell_total + 4095) / 4096 * 4096 : c.cell_total);
    k.cell_act_count = c.cell_act_count.p;
    k.frec_base = c.frec_base.p;
    k.worklist = c.worklist.p;
    k.worklist_sys = c.worklist_sys.p;
    k.prec = c.prec.p;
    k.brick_count = c.brick_count.p;
    k.brick_rec = c.brick_rec.p;
    k.brick_total = c.brick_total;
    k.brick_max = c.brick_max;
    k.any_bricks = c.any_bricks ? 1 : 0;
    k.any_cells = c.any_cells ? 1 : 0;
    k.frec0 = c.frec0.p;
    k.frec1 = c.frec1.p;
    k.frec2 = c.frec2.p;
    k.vrec = c.vrec.p;
    k.vval = c.vval.p;
    k.vsig = c.vsig.p;
    k.integrals = c.integrals.p;
    k.born = c.born.p;
    k.dbdi = c.dbdi.p;
    k.dedi = c.dedi.p;
    k.pair_part = c.pair_part.p;
    k.colpart = c.colpart.p;
    k.atom_energy = c.atom_energy.p;
    k.dedq = c.dedq.p;
    k.dedr_fix = c.dedr_fix.p;
    k.dedr = c.dedr.p;
    k.vacc_fix = c.vacc_fix.p;
    k.theta_atom = c.theta_atom.p;
    k.theta = c.theta.p;
    k.force_fix = c.force_fix.p;
    k.energy = c.energy.p;
    k.work = c.work.p;
    k.scan_tmp = c.scan_tmp.p;
    return k;
}

gbi_status ensure_adjoint_buffers(gbi_ctx& c) {
    if (c.theta.p) return GBI_OK;
    c.drop_graphs();
    HIP_TRY(hipStreamSynchronize(c.stream));
    HIP_TRY(c.vsig.ensure(c.NQ));
    HIP_TRY(c.dedq.ensure(c.N));
    HIP_TRY(c.dedr.ensure(c.N));
    HIP_TRY(c.dedr_fix.ensure(c.N));
    HIP_TRY(c.vacc_fix.ensure((size_t)c.S * gbi::kVaccCopies * gbi::kVaccStride));
    HIP_TRY(c.theta_atom.ensure((size_t)gbi::TH_COUNT * c.N));
    HIP_TRY(c.theta.ensure((size_t)c.S * gbi::TH_COUNT));
    return GBI_OK;
}

// Box edges a periodic system accepts: at least twice the cutoff (CharmmGBMVForceImpl.cpp:99-104,
// CharmmGBSWForceImpl.cpp:48-54), and more than twice a cell list's reach plus a cell diagonal,
// so a list holds one image of an atom at most and that image is the nearest to every point of
// its cell.
gbi_status check_box(const gbi_ctx& c, const SysHost& s, const double* L) {
    double maxR = 0.0;
    for (int j = s.atom_begin; j < s.atom_begin + s.N; ++j) maxR = std::max(maxR, c.radius[j]);
    const double reach = s.kind == gbi::GRID_FAITHFUL ? maxR + s.halfdiag + s.P.buffer : s.max_reach;
    for (int d = 0; d < 3; ++d) {
        if (!std::isfinite(L[d]) || !(L[d] > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "box edges must be positive");
        if (s.P.cutoff > 0.5 * L[d])
            return fail(GBI_ERR_INVALID_ARGUMENT,
                        "The cutoff distance cannot be greater than half the periodic box size.");
        if (!(L[d] > 2.0 * (reach + std::sqrt(3.0) * s.g)))
            return fail(GBI_ERR_INVALID_ARGUMENT, "periodic box smaller than twice the lookup-grid reach");
    }
    return GBI_OK;
}

gbi_status run(gbi_ctx& c, uint32_t flags) {
    if (!c.have_positions) return fail(GBI_ERR_INVALID_ARGUMENT, "positions were never set");
    const bool adjoint = (flags & GBI_ADJOINT) != 0;
    for (const SysHost& s : c.sys)
        if (s.periodic()) {
            if (!s.box_set)
                return fail(GBI_ERR_INVALID_ARGUMENT, "CutoffPeriodic system without a periodic box (gbi_set_periodic_box)");
            ST_TRY(check_box(c, s, s.box));
        }
    if (adjoint && c.cprog) return fail(GBI_ERR_UNSUPPORTED, "the adjoint covers the canonical GB forms only");
    if (c.cell_total == 0) ST_TRY(size_from_device(c, 0));
    if (adjoint) ST_TRY(ensure_adjoint_buffers(c));
    gbi::KParams k = make_kparams(c, adjoint);
    if (c.use_graph && !c.profiling) {
        hipGraphExec_t& ge = c.graph[adjoint ? 1 : 0];
        if (!ge) {
            hipGraph_t g = nullptr;
            HIP_TRY(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            const hipError_t le = gbi::launch_evaluate(k, c.kinds, c.stream, nullptr);
            const hipError_t ce = hipStreamEndCapture(c.stream, &g);
            HIP_TRY(le);
            HIP_TRY(ce);
            const hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) ge = nullptr;
            HIP_TRY(ie);
        }
        HIP_TRY(hipGraphLaunch(ge, c.stream));
    } else {
        const hipEvent_t* marks = nullptr;
        if (c.profiling && c.prof_used < 4096) {
            if (c.prof_used == c.prof_events.size()) {
                std::array<hipEvent_t, gbi::kPasses + 1> a;
                for (auto& e : a) HIP_TRY(hipEventCreate(&e));
                c.prof_events.push_back(a);
            }
            marks = c.prof_events[c.prof_used++].data();
        }
        HIP_TRY(gbi::launch_evaluate(k, c.kinds, c.stream, marks));
    }
    c.computed = true;
    c.computed_adjoint = adjoint;
    return GBI_OK;
}

// Device flags of the last evaluation: errors, or capacity growth and a redo.
gbi_status check_flags(gbi_ctx& c, bool& retry) {
    retry = false;
    std::vector<gbi::SysGeom> G(c.S);
    int all = 0, total_entries = 0;
    HIP_TRY(hipMemcpyAsync(G.data(), c.geom.p, sizeof(gbi::SysGeom) * c.S, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipMemcpyAsync(&all, c.flags_all.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipMemcpyAsync(&total_entries, c.cell_start.p + c.cell_total, sizeof(int), hipMemcpyDeviceToHost,
                           c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    for (int s = 0; s < c.S; ++s) all |= G[s].flags;
    if (all & gbi::FLAG_NONFINITE_POS) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    if (all & gbi::FLAG_GRID_TOO_LARGE)
        return fail(GBI_ERR_GRID_TOO_LARGE,
                    "CharmmGBMVForce: CustomGBIntegral lookup table dimension is too large, check atom positions!");
    if (all & (gbi::FLAG_CELL_OVERFLOW | gbi::FLAG_ENTRY_OVERFLOW | gbi::FLAG_POINT_OVERFLOW | gbi::FLAG_LIST_OVERFLOW)) {
        bool grew = false;
        if (all & gbi::FLAG_LIST_OVERFLOW) {
            // the fill's counts are exact past the capacity: room for the largest plus 25 %
            std::vector<int> fill(c.cell_total);
            HIP_TRY(hipMemcpyAsync(fill.data(), c.cell_fill.p, sizeof(int) * c.cell_total, hipMemcpyDeviceToHost,
                                   c.stream));
            HIP_TRY(hipStreamSynchronize(c.stream));
            long long cb = 0;
            for (int s = 0; s < c.S; ++s) {
                SysHost& h = c.sys[s];
                if (G[s].flags & gbi::FLAG_LIST_OVERFLOW) {
                    int mx = 0;
                    for (long long k = cb; k < cb + h.cell_cap && k < c.cell_total; ++k) mx = std::max(mx, fill[k]);
                    h.tmp_cap = std::max(h.tmp_cap + 1, mx + mx / 4 + 16);
                    grew = true;
                }
                cb += h.cell_cap;
            }
        }
        for (int s = 0; s < c.S; ++s) {
            SysHost& h = c.sys[s];
            if (G[s].flags & gbi::FLAG_CELL_OVERFLOW) {
                h.cell_cap *= 2;
                grew = true;
            }
            if (G[s].flags & gbi::FLAG_POINT_OVERFLOW) {
                h.pt_cap = (int)std::min<long long>(2ll * h.pt_cap, (long long)h.N * h.Q);
                grew = true;
            }
        }
        long long min_entries = 0;
        if (all & gbi::FLAG_ENTRY_OVERFLOW)
            min_entries = std::max<long long>(2ll * c.entry_cap, (long long)total_entries * 5 / 4);
        if (grew) ST_TRY(layout(c));
        ST_TRY(size_from_device(c, min_entries));
        retry = true;
        return GBI_OK;
    }
    if (all & gbi::FLAG_NONFINITE_RESULT) return fail(GBI_ERR_NONFINITE, "non-finite Born radius, energy or force");
    return GBI_OK;
}

gbi_status run_checked(gbi_ctx& c, uint32_t flags) {
    for (int attempt = 0; attempt < 6; ++attempt) {
        ST_TRY(run(c, flags));
        bool retry = false;
        ST_TRY(check_flags(c, retry));
        if (!retry) return GBI_OK;
    }
    return fail(GBI_ERR_CAPACITY, "lookup grid kept outgrowing its buffers");
}

gbi_status apply_tight_grid(gbi_ctx& c, double g) {
    bool changed = false;
    for (SysHost& s : c.sys)
        if (s.kind == gbi::GRID_TIGHT && s.g != g) {
            s.g = g;
            s.cell_cap = 0;
            s.tmp_cap = 0;
            const double est = 1.6 * 100.0 * s.Q * g * g * g + 64.0;
            s.pt_cap = std::max(1, (int)std::min<double>(est, (double)s.N * s.Q));
            changed = true;
        }
    if (!changed) return GBI_OK;
    c.cell_total = 0;  // re-sized from the positions on the next evaluation
    c.drop_graphs();
    return GBI_OK;
}

// The host description of a batch (systems, atoms, quadrature rows): no device work.
gbi_status setup_host(gbi_ctx& c, int S, const int32_t* n_atoms, const double* charge, const double* radius,
                      const gbi_params* params, bool custom) {
    if (S <= 0) return fail(GBI_ERR_INVALID_ARGUMENT, "n_systems must be positive");
    long long Ntot = 0, NQtot = 0;
    for (int s = 0; s < S; ++s) {
        ST_TRY(validate_params(params[s], n_atoms[s], custom));
        Ntot += n_atoms[s];
        // each system's points start on a multiple of 4 (16-byte vector loads of its V)
        NQtot += ((long long)n_atoms[s] * params[s].n_radial * params[s].n_lebedev + 3) & ~3ll;
    }
    if (Ntot >= gbi::kMaxAtoms) return fail(GBI_ERR_INVALID_ARGUMENT, "too many atoms (at most 2^21 - 1)");
    if (NQtot >= (1ll << 31)) return fail(GBI_ERR_INVALID_ARGUMENT, "N * Q must be below 2^31");
    ST_TRY(check_atoms((int)Ntot, charge, radius));
    c.S = S;
    c.N = (int)Ntot;
    c.NQ = (int)NQtot;
    c.charge.assign(charge, charge + Ntot);
    c.radius.assign(radius, radius + Ntot);
    c.tuning.blocks_per_cu = kDefaultBlocksPerCU;
    c.tuning.pair_columns = 0;
    c.tuning.tight_grid_length = 0.0;
    int atom = 0, qrow = 0, pt = 0;
    for (int s = 0; s < S; ++s) {
        SysHost h;
        h.P = params[s];
        if (h.P.model == GBI_MODEL_GBSW) h.P.integral_type = GBI_INTEGRAL_GBSW;
        if (h.P.tight_grid_length <= 0.0) h.P.tight_grid_length = kDefaultTightGrid;
        h.N = n_atoms[s];
        h.atom_begin = atom;
        h.Q = h.P.n_radial * h.P.n_lebedev;
        h.quad_base = qrow;
        h.pt_base = pt;
        h.vol = (gbi::VolKind)h.P.integral_type;
        if (h.P.lookup_policy == GBI_LOOKUP_BRUTE)
            h.kind = gbi::GRID_BRUTE;
        else if (h.P.model == GBI_MODEL_GBSW)
            h.kind = gbi::GRID_TIGHT;  // the standalone lists are complete, so exact lists give the same V
        else
            h.kind = gbi::GRID_FAITHFUL;
        h.g = h.kind == gbi::GRID_FAITHFUL ? h.P.grid_length : h.P.tight_grid_length;
        h.halfdiag = std::sqrt(3.0) / 2.0 * h.P.grid_length;
        h.gpad = 0.03 + h.halfdiag + h.P.buffer;  // CustomGBIntegral.cpp:174-175
        const double est = 1.6 * 100.0 * h.Q * h.g * h.g * h.g + 64.0;
        h.pt_cap = h.kind == gbi::GRID_BRUTE ? h.N * h.Q : (int)std::min<double>(est, (double)h.N * h.Q);
        h.pt_cap = std::max(h.pt_cap, 1);
        c.kinds[h.vol] = true;
        c.any_cut = c.any_cut || h.cutoff();
        c.any_trunc = c.any_trunc || (h.cutoff() && h.P.model == GBI_MODEL_GBSW);
        atom += h.N;
        qrow += h.Q;
        pt += (h.N * h.Q + 3) & ~3;  // 4-aligned point blocks (NQ counts the padding)
        c.sys.push_back(h);
    }
    c.quad_rows = qrow;
    return GBI_OK;
}

// Host tables of a batch: quadrature offsets, w_q / r_q^(order-2) per integral and |off_q|^2
// (one block of rows per system), the system of every atom, and the binning order (systems in
// turn, each with a stride coprime to its size so that concurrent workgroups hit distant
// cells of the same system's counters).
struct HostTables {
    std::vector<double> off, pw;
    std::vector<float> r2;
    std::vector<int> atom_sys, bin_order, pair_atom;
};

gbi_status host_tables(const gbi_ctx& c, HostTables& t) {
    const int qrow = c.quad_rows;
    t.off.assign(3 * (size_t)qrow, 0.0);
    t.pw.assign(2 * (size_t)qrow, 0.0);
    t.r2.assign(qrow, 0.f);
    for (const SysHost& h : c.sys) {
        std::vector<std::array<double, 5>> table;
        try {
            table = gbi::quadrature_table(h.P.n_radial, h.P.r0, h.P.r1, h.P.n_lebedev);
        } catch (const std::exception& e) {
            return fail(GBI_ERR_INVALID_ARGUMENT, e.what());
        }
        for (int q = 0; q < h.Q; ++q) {
            const size_t row = (size_t)h.quad_base + q;
            for (int d = 0; d < 3; ++d) t.off[3 * row + d] = table[q][d];
            for (int k = 0; k < 2; ++k)
                t.pw[k * (size_t)qrow + row] = table[q][4] / std::pow(table[q][3], h.P.orders[k] - 2);
            t.r2[row] = (float)(table[q][3] * table[q][3]);
        }
    }
    t.atom_sys.assign(c.N, 0);
    for (int s = 0; s < c.S; ++s)
        std::fill(t.atom_sys.begin() + c.sys[s].atom_begin, t.atom_sys.begin() + c.sys[s].atom_begin + c.sys[s].N, s);
    // binning order: each system's atoms with a stride coprime to its size, the systems
    // interleaved round-robin, so the workgroups in flight at once spread over every system's
    // counters.  System after system, most of them belonged to one small system and contended
    // on its counters: 8 conformations of 2,000 atoms took 0.41 ms to bin, about what one
    // 25,000-atom system takes; interleaved, the adjoint of that batch runs at 8,550 against
    // 7,309 conformation evaluations per second, the 64-system mix at 4,294 against 4,120.
    t.bin_order.assign(c.N, 0);
    std::vector<std::vector<int>> seq(c.S);
    for (int s = 0; s < c.S; ++s) {
        const SysHost& h = c.sys[s];
        long long stride = 1;
        for (int cand : {7919, 7927, 7933, 7937, 7949})
            if (h.N % cand != 0) {
                stride = cand;
                break;
            }
        seq[s].resize(h.N);
        for (int k = 0; k < h.N; ++k) seq[s][k] = h.atom_begin + (int)((k * stride) % h.N);
    }
    std::vector<int> next(c.S, 0);
    for (int w = 0; w < c.N;)
        for (int s = 0; s < c.S; ++s)
            if (next[s] < c.sys[s].N) t.bin_order[w++] = seq[s][next[s]++];
    t.pair_atom.resize(c.N);
    for (int k = 0; k < c.N; ++k) t.pair_atom[k] = k;
    return GBI_OK;
}

gbi_status create_batch(int S, const int32_t* n_atoms, const double* charge, const double* radius,
                        const gbi_params* params, gbi_ctx** out, bool custom = false) {
    if (!out || !n_atoms || !charge || !radius || !params) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    auto owned = std::make_unique<gbi_ctx>();
    gbi_ctx* c = owned.get();
    ST_TRY(setup_host(*c, S, n_atoms, charge, radius, params, custom));
    HostTables t;
    ST_TRY(host_tables(*c, t));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GBI_ERR_NO_DEVICE, "no HIP device");
    (void)hipGetDevice(&c->device);
    if (hipDeviceGetAttribute(&c->num_cus, hipDeviceAttributeMultiprocessorCount, c->device) != hipSuccess ||
        c->num_cus <= 0)
        c->num_cus = 256;
    HIP_TRY(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    const size_t N = c->N, NQ = c->NQ, qrow = c->quad_rows;
    HIP_TRY(c->d_sys.ensure(S));
    HIP_TRY(c->geom.ensure(S));
    HIP_TRY(c->d_box.ensure(3 * (size_t)S));
    HIP_TRY(hipMemsetAsync(c->d_box.p, 0, sizeof(double) * 3 * S, c->stream));
    HIP_TRY(c->flags_all.ensure(1));
    HIP_TRY(c->bbox_part.ensure((size_t)S * gbi::kBboxMaxSlices * gbi::kBboxPart));
    HIP_TRY(c->bbox_done.ensure(S));
    HIP_TRY(c->atom_sys.ensure(N));
    HIP_TRY(c->pos.ensure(3 * N));
    HIP_TRY(c->d_charge.ensure(N));
    HIP_TRY(c->d_radius.ensure(N));
    HIP_TRY(c->d_reach.ensure(N));
    HIP_TRY(c->aparam.ensure(N));
    HIP_TRY(c->quad_off.ensure(3 * qrow));
    HIP_TRY(c->quad_pw.ensure(2 * qrow));
    HIP_TRY(c->quad_r2.ensure(qrow));
    HIP_TRY(c->integrals.ensure(2 * N));
    HIP_TRY(c->born.ensure(N));
    HIP_TRY(c->dbdi.ensure(2 * N));
    HIP_TRY(c->dedi.ensure(2 * N));
    HIP_TRY(c->atom_energy.ensure(N));
    HIP_TRY(c->energy.ensure(S));
    HIP_TRY(c->stats.ensure(6));
    HIP_TRY(c->work.ensure(8));
    HIP_TRY(c->frec0.ensure(NQ));
    HIP_TRY(c->frec1.ensure(NQ));
    HIP_TRY(c->frec2.ensure(c->kinds[gbi::VOL_GBMV2] ? NQ : 1));
    HIP_TRY(c->vrec.ensure(NQ));
    HIP_TRY(c->vval.ensure(NQ));
    HIP_TRY(c->force_fix.ensure(3 * N));
    HIP_TRY(c->bin_order.ensure(N));
    HIP_TRY(c->pair_atom.ensure(N));
    HIP_TRY(hipMemcpyAsync(c->quad_off.p, t.off.data(), sizeof(double) * t.off.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->quad_pw.p, t.pw.data(), sizeof(double) * t.pw.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->quad_r2.p, t.r2.data(), sizeof(float) * t.r2.size(), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->atom_sys.p, t.atom_sys.data(), sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->pair_atom.p, t.pair_atom.data(), sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->bin_order.p, t.bin_order.data(), sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->pair_order_host = t.pair_atom;
    ST_TRY(upload_atoms(*c));
    ST_TRY(build_tiles(*c));
    *out = owned.release();
    return GBI_OK;
}

// ---------------------------------------------------------------------------
// gbi_execute_noalloc: one system evaluated without any device allocation.  Every device
// buffer is a piece of the library's static scratch array (gbi_noalloc.hip), its capacities
// computed on the host from the positions: the lookup grid the device will build (the same
// FP64 expressions as bbox_to_geom, candidate_box, cell_member and point_cell), each cell's
// raw members and quadrature points, with margins for decisions that could round the other
// way at a cell face.  A grid that still outgrows them is an error, not a regrowth.
// ---------------------------------------------------------------------------

struct NoAllocCaps {
    long long cells = 0, entries = 0;
    int tmp_cap = 0, pt_cap = 0;
};

NoAllocCaps noalloc_caps(const SysHost& s, const double* pos, const double* radius, const double* reach) {
    NoAllocCaps caps;
    const int N = s.N;
    if (s.kind == gbi::GRID_BRUTE) {
        caps.cells = 1;
        caps.entries = N;
        caps.tmp_cap = N;
        caps.pt_cap = N * s.Q;
        return caps;
    }
    const bool faithful = s.kind == gbi::GRID_FAITHFUL;
    const double g = s.g, inv_g = 1.0 / s.g;
    const double gpad = faithful ? s.gpad : s.max_reach;
    double mn[3], mx[3], origin[3], gmax[3];
    int n[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = pos[d];
    for (int a = 0; a < N; ++a)
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], pos[3 * a + d]);
            mx[d] = std::max(mx[d], pos[3 * a + d]);
        }
    for (int d = 0; d < 3; ++d) {
        const double lo = mn[d] - gpad, hi = mx[d] + gpad, len = hi - lo;
        n[d] = len > 1000.0 ? 1 : (faithful ? (int)std::ceil(len / g) + 1 : (int)std::floor(len * inv_g) + 1);
        origin[d] = lo;
        gmax[d] = faithful ? lo + (n[d] - 1) * g : lo + n[d] * g;
    }
    const long long ncell = (long long)n[0] * n[1] * n[2];
    caps.cells = (long long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);  // room for one more cell per axis
    std::vector<int> raw(ncell, 0), pts(ncell, 0);
    for (int j = 0; j < N; ++j) {
        const double* x = pos + 3 * j;
        const double pad = faithful ? (radius[j] + s.halfdiag) + s.P.buffer : reach[j];
        int b[3], e[3];
        for (int d = 0; d < 3; ++d) {
            b[d] = faithful ? (int)std::floor((x[d] - pad - origin[d]) / g) : (int)std::floor((x[d] - pad - origin[d]) * inv_g);
            e[d] = faithful ? (int)std::ceil((x[d] + pad - origin[d]) / g) : (int)std::floor((x[d] + pad - origin[d]) * inv_g);
            b[d] = std::max(b[d], 0);
            e[d] = std::min(e[d], n[d] - 1);
        }
        for (int i = b[0]; i <= e[0]; ++i)
            for (int jj = b[1]; jj <= e[1]; ++jj)
                for (int k = b[2]; k <= e[2]; ++k) {
                    const double lx = origin[0] + i * g, ly = origin[1] + jj * g, lz = origin[2] + k * g;
                    double dx, dy, dz;
                    if (faithful) {
                        dx = lx - x[0];
                        dy = ly - x[1];
                        dz = lz - x[2];
                    } else {
                        dx = std::max(std::max(lx - x[0], x[0] - (lx + g)), 0.0);
                        dy = std::max(std::max(ly - x[1], x[1] - (ly + g)), 0.0);
                        dz = std::max(std::max(lz - x[2], x[2] - (lz + g)), 0.0);
                    }
                    if (dx * dx + dy * dy + dz * dz < pad * pad) ++raw[((long long)i * n[1] + jj) * n[2] + k];
                }
    }
    std::vector<std::array<double, 5>> table = gbi::quadrature_table(s.P.n_radial, s.P.r0, s.P.r1, s.P.n_lebedev);
    for (int i = 0; i < N; ++i)
        for (int q = 0; q < s.Q; ++q) {
            long long c = 0;
            for (int d = 0; d < 3; ++d) {
                const double pt = pos[3 * i + d] + table[q][d];
                int idx;
                if (faithful) {
                    if (pt < origin[d] || pt > gmax[d]) c = -1;
                    idx = (int)std::floor((pt - origin[d]) / g);
                } else {
                    const double t = (pt - origin[d]) * inv_g;
                    if (!(t >= 0.0) || t >= (double)n[d]) c = -1;
                    idx = (int)std::floor(t);
                }
                if (c >= 0) c = c * n[d] + std::min(std::max(idx, 0), n[d] - 1);
            }
            if (c >= 0 && raw[c] > 0) ++pts[c];
        }
    int max_raw = 0, max_pts = 0;
    long long members = 0;
    const int cap = (s.P.cell_cap > 0) ? s.P.cell_cap : INT_MAX;
    for (long long c = 0; c < ncell; ++c) {
        max_raw = std::max(max_raw, raw[c]);
        max_pts = std::max(max_pts, pts[c]);
        members += std::min(raw[c], cap);
    }
    caps.tmp_cap = max_raw + max_raw / 8 + 16;
    caps.pt_cap = max_pts + max_pts / 8 + 16;
    caps.entries = members + members / 8 + 1024;
    return caps;
}

// Pieces of the scratch array (256-byte aligned) handed to buffers of a context, which never
// frees them; without a base it only adds up the bytes.
struct ScratchCarver {
    char* base = nullptr;
    size_t used = 0;
    template <class T>
    void take(DevBuf<T>& b, size_t count) {
        count = std::max<size_t>(count, 1);
        used = (used + 255) & ~size_t(255);
        if (base) {
            b.p = reinterpret_cast<T*>(base + used);
            b.n = count;
        }
        used += count * sizeof(T);
    }
};

}  // namespace

extern "C" {

int32_t gbi_abi_version(void) { return GBI_ABI_VERSION; }

const char* gbi_last_error(void) { return g_last_error.c_str(); }

gbi_status gbi_params_default(int32_t model, int32_t integral_type, gbi_params* out) {
    if (!out) return fail(GBI_ERR_INVALID_ARGUMENT, "null output");
    gbi_params P;
    std::memset(&P, 0, sizeof(P));
    P.model = model;
    P.integral_type = model == GBI_MODEL_GBSW ? GBI_INTEGRAL_GBSW : integral_type;
    P.lookup_policy = GBI_LOOKUP_REFERENCE;
    P.nonbonded_method = GBI_NO_CUTOFF;
    P.cutoff = 1.0;
    P.n_radial = 24;
    P.r0 = 0.05;
    P.r1 = 2.0;
    P.grid_length = 0.15;
    P.orders[0] = 4;
    P.orders[1] = 7;
    P.sw = 0.03;
    P.rscale = 0.952;
    P.roffset = 0.03;
    P.vguard = 1e-30;
    P.alpha0 = -0.180;
    P.alpha1 = 1.817;
    P.eta = 0.05;
    P.slope = 1.0;
    P.shift = 0.0;
    P.prefactor = -kOne4PiEps0 * (1.0 - 1.0 / 80.0);  // CharmmReferenceGBSW.cpp:177-178
    P.tight_grid_length = kDefaultTightGrid;
    if (model == GBI_MODEL_GBSW) {
        // CharmmReferenceGBSW.cpp:20-52: rule 4 (38 points), buffer 0.20 nm, no cap
        P.n_lebedev = 38;
        P.buffer = 0.20;
        P.cell_cap = 0;
    } else if (model == GBI_MODEL_GBMV) {
        P.n_lebedev = 26;  // rule 3 (GBSWIntegral.cpp:40, GBMVIntegralType{I,II}.cpp:36,40)
        P.cell_cap = 64;   // CustomGBIntegral.cpp:15
        if (integral_type == GBI_INTEGRAL_GBSW) {
            P.buffer = 0.01 + P.sw;  // GBSWIntegral.cpp:60
        } else if (integral_type == GBI_INTEGRAL_GBMV1) {
            P.buffer = 0.21;  // GBMVIntegralTypeI.cpp:22-24,56
            P.gamma0 = 0.44;
            P.beta = -100.0;
            P.lambda = 0.1;
        } else if (integral_type == GBI_INTEGRAL_GBMV2) {
            P.buffer = 0.21;  // GBMVIntegralTypeII.cpp:23-28,60
            P.gamma0 = 0.44;
            P.beta = -20.0;
            P.lambda = 0.5;
            P.P1 = 0.45;
            P.P2 = 1.25;
            P.S0 = 0.7;
        } else if (integral_type == GBI_INTEGRAL_CUSTOM) {
            *out = P;
            return fail(GBI_ERR_UNSUPPORTED, "CustomIntegral (type 3) is not implemented by the reference either");
        } else {
            return fail(GBI_ERR_INVALID_ARGUMENT, "unknown GB integral type");
        }
    } else {
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown model");
    }
    *out = P;
    return GBI_OK;
}

gbi_status gbi_quadrature(int32_t n_radial, double r0, double r1, int32_t n_lebedev, double* out) {
    if (!out) return fail(GBI_ERR_INVALID_ARGUMENT, "null output");
    try {
        const auto t = gbi::quadrature_table(n_radial, r0, r1, n_lebedev);
        for (size_t q = 0; q < t.size(); ++q)
            for (int k = 0; k < 5; ++k) out[5 * q + k] = t[q][k];
    } catch (const std::exception& e) {
        return fail(GBI_ERR_INVALID_ARGUMENT, e.what());
    }
    return GBI_OK;
}

gbi_status gbi_create(int32_t n_atoms, const double* charge, const double* radius, const gbi_params* params,
                      gbi_ctx** out) {
    return create_batch(1, &n_atoms, charge, radius, params, out);
}

gbi_status gbi_batch_create(int32_t n_systems, const int32_t* n_atoms, const double* charge, const double* radius,
                            const gbi_params* params, gbi_ctx** out) {
    return create_batch(n_systems, n_atoms, charge, radius, params, out);
}

gbi_status gbi_get_layout(gbi_ctx* ctx, int32_t* n_systems, int32_t* n_atoms_total) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    if (n_systems) *n_systems = ctx->S;
    if (n_atoms_total) *n_atoms_total = ctx->N;
    return GBI_OK;
}

void gbi_destroy(gbi_ctx* ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
}

gbi_status gbi_set_positions_posq(gbi_ctx* ctx, const void* posq, int32_t posq_double, const int32_t* atom_slot) {
    if (!ctx || !posq) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(gbi::launch_gather_posq(ctx->pos.p, posq, posq_double ? 1 : 0, atom_slot, ctx->N, ctx->stream));
    ctx->have_positions = true;
    return GBI_OK;
}

gbi_status gbi_add_forces_fixed(gbi_ctx* ctx, int64_t* force, int32_t padded_n, const int32_t* atom_slot) {
    if (!ctx || !force) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (padded_n < ctx->N) return fail(GBI_ERR_INVALID_ARGUMENT, "padded_n is below the atom count");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "nothing computed yet");
    HIP_TRY(gbi::launch_add_forces_fixed(reinterpret_cast<long long*>(force), padded_n, ctx->force_fix.p, atom_slot,
                                         ctx->N, ctx->stream));
    return GBI_OK;
}

gbi_status gbi_set_stream(gbi_ctx* ctx, void* hip_stream) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return GBI_OK;
}

gbi_status gbi_update_parameters(gbi_ctx* ctx, int32_t n_atoms, const double* charge, const double* radius) {
    if (!ctx || !charge || !radius) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (n_atoms != ctx->N)  // CharmmReferenceKernels.cpp:373-374
        return fail(GBI_ERR_PARTICLE_COUNT, "updateParametersInContext: The number of particles has changed");
    if (ctx->cprog)
        return fail(GBI_ERR_UNSUPPORTED, "a general-expression context takes its parameters at gbi_create_custom");
    ST_TRY(check_atoms(n_atoms, charge, radius));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->charge.assign(charge, charge + n_atoms);
    ctx->radius.assign(radius, radius + n_atoms);
    ST_TRY(upload_atoms(*ctx));
    // radii change the lookup reach: re-estimate the grid
    if (ctx->have_positions) {
        if (ctx->cell_total > 0) ST_TRY(layout(*ctx));  // new reach in the descriptors
        return size_from_device(*ctx, 0);
    }
    return GBI_OK;
}

gbi_status gbi_set_periodic_box(gbi_ctx* ctx, int32_t system, const double* v) {
    if (!ctx || !v) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (system < 0 || system >= ctx->S) return fail(GBI_ERR_INVALID_ARGUMENT, "no such system");
    SysHost& s = ctx->sys[system];
    if (!s.periodic()) return fail(GBI_ERR_INVALID_ARGUMENT, "the system is not CutoffPeriodic");
    if (v[1] != 0.0 || v[2] != 0.0 || v[3] != 0.0 || v[5] != 0.0 || v[6] != 0.0 || v[7] != 0.0)
        return fail(GBI_ERR_UNSUPPORTED, "CharmmGBMVForce: Only Rectangular Periodic Box is Supported!");
    const double L[3] = {v[0], v[4], v[8]};
    ST_TRY(check_box(*ctx, s, L));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int d = 0; d < 3; ++d) s.box[d] = L[d];
    s.box_set = true;
    HIP_TRY(hipMemcpyAsync(ctx->d_box.p + 3 * system, s.box, sizeof(double) * 3, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_set_exclusions(gbi_ctx* ctx, int32_t system, int32_t n_pairs, const int32_t* pairs) {
    if (!ctx || (n_pairs > 0 && !pairs)) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (system < 0 || system >= ctx->S) return fail(GBI_ERR_INVALID_ARGUMENT, "no such system");
    if (n_pairs < 0) return fail(GBI_ERR_INVALID_ARGUMENT, "negative pair count");
    SysHost& s = ctx->sys[system];
    if (s.P.model != GBI_MODEL_GBMV && n_pairs > 0)
        return fail(GBI_ERR_INVALID_ARGUMENT, "CharmmGBSWForce has no exclusions");
    for (int32_t k = 0; k < 2 * n_pairs; k += 2)
        if (pairs[k] < 0 || pairs[k + 1] < 0 || pairs[k] >= s.N || pairs[k + 1] >= s.N || pairs[k] == pairs[k + 1])
            return fail(GBI_ERR_INVALID_ARGUMENT, "excluded pair out of range or of an atom with itself");
    s.excl.assign(pairs, pairs + 2 * (size_t)n_pairs);
    if (ctx->cprog) return build_custom_exclusions(*ctx);
    return build_exclusions(*ctx);
}

gbi_status gbi_create_custom(int32_t n_atoms, const gbi_params* params, const gbi_custom_spec* spec, gbi_ctx** out) {
    if (!params || !spec || !out) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (params->model != GBI_MODEL_GBMV) return fail(GBI_ERR_INVALID_ARGUMENT, "general expressions need GBI_MODEL_GBMV");
    if (n_atoms <= 0) return fail(GBI_ERR_INVALID_ARGUMENT, "n_atoms must be positive");
    if (spec->n_params < 1 || !spec->param_names || !spec->param_values || spec->radius_param < 0 ||
        spec->radius_param >= spec->n_params)
        return fail(GBI_ERR_INVALID_ARGUMENT, "GBSWIntegral: the per perticle parameter 'radius' must be defined");
    if (spec->n_integrals < 1 || spec->n_integrals > gbi::kMaxCustomIntegrals || !spec->integral_names)
        return fail(GBI_ERR_UNSUPPORTED, "general expressions take one or two integrals");
    if (spec->n_values < 0 || spec->n_values > gbi::kMaxCustomValues || spec->n_terms < 1 ||
        (spec->n_values > 0 && (!spec->value_names || !spec->value_exprs)) || !spec->term_exprs || !spec->term_types ||
        (spec->n_globals > 0 && (!spec->global_names || !spec->global_values)) || spec->n_globals < 0)
        return fail(GBI_ERR_INVALID_ARGUMENT, "incomplete expression specification");
    gbi::CustomSpec cs;
    for (int k = 0; k < spec->n_params; ++k) cs.params.push_back(spec->param_names[k] ? spec->param_names[k] : "");
    for (int k = 0; k < spec->n_globals; ++k) cs.globals.push_back(spec->global_names[k] ? spec->global_names[k] : "");
    for (int k = 0; k < spec->n_integrals; ++k)
        cs.integrals.push_back(spec->integral_names[k] ? spec->integral_names[k] : "");
    for (int k = 0; k < spec->n_values; ++k) {
        cs.value_names.push_back(spec->value_names[k] ? spec->value_names[k] : "");
        cs.value_exprs.push_back(spec->value_exprs[k] ? spec->value_exprs[k] : "");
    }
    for (int k = 0; k < spec->n_terms; ++k) {
        cs.term_exprs.push_back(spec->term_exprs[k] ? spec->term_exprs[k] : "");
        cs.term_types.push_back(spec->term_types[k]);
    }
    if (spec->n_param_derivs < 0 || spec->n_param_derivs > gbi::kMaxCustomDerivs ||
        (spec->n_param_derivs > 0 && !spec->param_deriv_names))
        return fail(GBI_ERR_INVALID_ARGUMENT, "at most 4 energy parameter derivatives, by name");
    for (int m = 0; m < spec->n_param_derivs; ++m) {
        const std::string name = spec->param_deriv_names[m] ? spec->param_deriv_names[m] : "";
        int g = -1;
        for (int k = 0; k < (int)cs.globals.size(); ++k)
            if (cs.globals[k] == name) g = k;
        if (g < 0)
            return fail(GBI_ERR_INVALID_ARGUMENT,
                        "energy parameter derivative of '" + name + "', which is not a global parameter");
        cs.deriv_globals.push_back(g);
    }
    if (spec->n_functions < 0 || (spec->n_functions > 0 && (!spec->function_names || !spec->function_kinds ||
                                                            !spec->function_sizes || !spec->function_periodic ||
                                                            !spec->function_ranges || !spec->function_values)))
        return fail(GBI_ERR_INVALID_ARGUMENT, "incomplete tabulated function specification");
    for (int k = 0, off = 0; k < spec->n_functions; ++k) {
        gbi::TabFunction f;
        f.name = spec->function_names[k] ? spec->function_names[k] : "";
        f.kind = spec->function_kinds[k];
        if (f.kind < gbi::TAB_CONTINUOUS_1D || f.kind > gbi::TAB_CONTINUOUS_3D)
            return fail(GBI_ERR_UNSUPPORTED, "tabulated function '" + f.name + "': unknown kind " +
                                                 std::to_string(f.kind));
        if (gbi::tab_arity(f.kind) > 1) {
            if (!spec->function_dims) return fail(GBI_ERR_INVALID_ARGUMENT, "multi-dimensional table without sizes");
            for (int d = 0; d < 3; ++d) f.dims[d] = spec->function_dims[3 * k + d];
        }
        if (f.kind == gbi::TAB_CONTINUOUS_2D || f.kind == gbi::TAB_CONTINUOUS_3D) {
            if (!spec->function_ranges_yz)
                return fail(GBI_ERR_INVALID_ARGUMENT, "Continuous2D/3D table without y and z ranges");
            f.ymin = spec->function_ranges_yz[4 * k];
            f.ymax = spec->function_ranges_yz[4 * k + 1];
            f.zmin = spec->function_ranges_yz[4 * k + 2];
            f.zmax = spec->function_ranges_yz[4 * k + 3];
        }
        const int n = spec->function_sizes[k];
        if (n < (f.kind == gbi::TAB_CONTINUOUS_1D ? 2 : 1))
            return fail(GBI_ERR_INVALID_ARGUMENT, "tabulated function '" + f.name + "' has too few values");
        f.periodic = spec->function_periodic[k] != 0;
        f.min = spec->function_ranges[2 * k];
        f.max = spec->function_ranges[2 * k + 1];
        if ((f.kind == gbi::TAB_CONTINUOUS_1D || f.kind == gbi::TAB_CONTINUOUS_2D || f.kind == gbi::TAB_CONTINUOUS_3D) &&
            !(f.min < f.max))
            return fail(GBI_ERR_INVALID_ARGUMENT, "tabulated function '" + f.name + "' needs min < max");
        f.values.assign(spec->function_values + off, spec->function_values + off + n);
        off += n;
        cs.functions.push_back(std::move(f));
    }
    // the expressions are checked before any device work
    auto prog = std::make_unique<gbi::CustomProgram>();
    try {
        gbi::generate_custom_source(cs, *prog);
    } catch (const std::exception& e) {
        return fail(GBI_ERR_INVALID_ARGUMENT, e.what());
    }
    gbi_params P = *params;
    if (spec->n_integrals == 1) P.orders[1] = P.orders[0];
    std::vector<double> radius(n_atoms), charge(n_atoms, 0.0);
    for (int a = 0; a < n_atoms; ++a) radius[a] = spec->param_values[(size_t)a * spec->n_params + spec->radius_param];
    for (size_t k = 0; k < (size_t)n_atoms * spec->n_params; ++k)
        if (!std::isfinite(spec->param_values[k])) return fail(GBI_ERR_NONFINITE, "non-finite per-particle parameter");
    gbi_ctx* c = nullptr;
    const int32_t n = n_atoms;
    ST_TRY(create_batch(1, &n, charge.data(), radius.data(), &P, &c, true));
    std::unique_ptr<gbi_ctx> guard(c);
    c->sys[0].custom = true;
    try {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) return fail(GBI_ERR_HIP, "hipGetDeviceProperties");
        gbi::compile_custom_program(prop.gcnArchName, *prog);
    } catch (const std::exception& e) {
        return fail(GBI_ERR_INVALID_ARGUMENT, e.what());
    }
    const size_t N = n_atoms;
    const int NV = prog->n_values, NPL = prog->n_planes;
    const size_t NDD = std::max(1, prog->n_derivs);
    HIP_TRY(c->c_prm.ensure(N * spec->n_params));
    HIP_TRY(c->c_glob.ensure(std::max(1, spec->n_globals)));
    HIP_TRY(c->c_vals.ensure(std::max<size_t>(1, NV * N)));
    HIP_TRY(c->c_dvdi.ensure(std::max<size_t>(1, 2 * NV * N)));
    HIP_TRY(c->c_esp.ensure(N));
    HIP_TRY(c->c_despdi.ensure(2 * N));
    HIP_TRY(c->c_dvdth.ensure(std::max<size_t>(1, NV) * NDD * N));
    HIP_TRY(c->c_despdth.ensure(NDD * N));
    HIP_TRY(c->c_pdth.ensure(NDD * N));
    HIP_TRY(c->c_dvdx.ensure(std::max<size_t>(1, NV) * 3 * N));
    HIP_TRY(c->c_despdx.ensure(3 * N));
    const size_t nb = (N + 63) / 64;
    HIP_TRY(c->c_part.ensure(nb * (nb + 1) / 2 * NPL * 64));
    HIP_TRY(hipMemcpyAsync(c->c_prm.p, spec->param_values, sizeof(double) * N * spec->n_params, hipMemcpyHostToDevice,
                           c->stream));
    if (spec->n_globals > 0)
        HIP_TRY(hipMemcpyAsync(c->c_glob.p, spec->global_values, sizeof(double) * spec->n_globals,
                               hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    gbi::CustomLaunch& L = c->claunch;
    L.values = prog->values;
    L.pairs = prog->pairs;
    L.pair_threads = 64 * gbi::kCustomPairWaves;
    L.chain_threads = 64 * gbi::kCustomChainWaves;
    L.chain = prog->chain;
    L.N = n_atoms;
    L.nb = (int)nb;
    L.method = P.nonbonded_method;
    L.cut2 = P.nonbonded_method != GBI_NO_CUTOFF ? P.cutoff * P.cutoff : 0.0;
    L.box = c->d_box.p;
    L.prm = c->c_prm.p;
    L.glob = c->c_glob.p;
    L.vals = c->c_vals.p;
    L.dvdi = c->c_dvdi.p;
    L.esp = c->c_esp.p;
    L.despdi = c->c_despdi.p;
    L.part = c->c_part.p;
    L.dvdth = c->c_dvdth.p;
    L.despdth = c->c_despdth.p;
    L.pdth = c->c_pdth.p;
    L.dvdx = c->c_dvdx.p;
    L.despdx = c->c_despdx.p;
    c->custom_nparams = spec->n_params;
    c->custom_radius = spec->radius_param;
    c->cprog = std::move(prog);
    ST_TRY(build_custom_exclusions(*c));
    *out = guard.release();
    return GBI_OK;
}

gbi_status gbi_set_custom_globals(gbi_ctx* ctx, int32_t n_globals, const double* values) {
    if (!ctx || (n_globals > 0 && !values)) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->cprog) return fail(GBI_ERR_INVALID_ARGUMENT, "not a general-expression context");
    if (n_globals != ctx->cprog->n_globals) return fail(GBI_ERR_INVALID_ARGUMENT, "wrong number of global parameters");
    if (n_globals == 0) return GBI_OK;
    HIP_TRY(hipMemcpyAsync(ctx->c_glob.p, values, sizeof(double) * n_globals, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_get_custom_param_derivs(gbi_ctx* ctx, int32_t n_derivs, double* out) {
    if (!ctx || (n_derivs > 0 && !out)) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->cprog) return fail(GBI_ERR_INVALID_ARGUMENT, "not a general-expression context");
    if (n_derivs != ctx->cprog->n_derivs) return fail(GBI_ERR_INVALID_ARGUMENT, "wrong number of parameter derivatives");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "no evaluation yet");
    if (n_derivs == 0) return GBI_OK;
    std::vector<double> per_atom((size_t)n_derivs * ctx->N);
    HIP_TRY(hipMemcpyAsync(per_atom.data(), ctx->c_pdth.p, sizeof(double) * per_atom.size(), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int m = 0; m < n_derivs; ++m) {  // fixed order: atom by atom
        double acc = 0.0;
        for (int a = 0; a < ctx->N; ++a) acc += per_atom[(size_t)m * ctx->N + a];
        out[m] = acc;
    }
    return GBI_OK;
}

gbi_status gbi_update_custom_parameters(gbi_ctx* ctx, int32_t n_atoms, const double* param_values) {
    if (!ctx || !param_values) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->cprog) return fail(GBI_ERR_INVALID_ARGUMENT, "not a general-expression context");
    if (n_atoms != ctx->N)  // CharmmReferenceKernels.cpp:373-374
        return fail(GBI_ERR_PARTICLE_COUNT, "updateParametersInContext: The number of particles has changed");
    const size_t np = ctx->custom_nparams;
    for (size_t k = 0; k < (size_t)n_atoms * np; ++k)
        if (!std::isfinite(param_values[k])) return fail(GBI_ERR_NONFINITE, "non-finite per-particle parameter");
    std::vector<double> radius(n_atoms);
    for (int a = 0; a < n_atoms; ++a) radius[a] = param_values[(size_t)a * np + ctx->custom_radius];
    ST_TRY(check_atoms(n_atoms, ctx->charge.data(), radius.data()));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->c_prm.p, param_values, sizeof(double) * n_atoms * np, hipMemcpyHostToDevice,
                           ctx->stream));
    ctx->radius = radius;
    ST_TRY(upload_atoms(*ctx));
    if (ctx->have_positions) {
        if (ctx->cell_total > 0) ST_TRY(layout(*ctx));
        return size_from_device(*ctx, 0);
    }
    return GBI_OK;
}

gbi_status gbi_update_constants(gbi_ctx* ctx, int32_t system, const gbi_params* params) {
    if (!ctx || !params) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (system < 0 || system >= ctx->S) return fail(GBI_ERR_INVALID_ARGUMENT, "no such system");
    SysHost& s = ctx->sys[system];
    gbi_params P = *params;
    if (P.model == GBI_MODEL_GBSW) P.integral_type = GBI_INTEGRAL_GBSW;
    if (P.tight_grid_length <= 0.0) P.tight_grid_length = s.P.tight_grid_length;
    if (s.custom) {  // a general-expression system keeps the integral orders of its spec
        P.orders[0] = s.P.orders[0];
        P.orders[1] = s.P.orders[1];
    }
    // everything but the constants must stay as it is
    const gbi_params& o = s.P;
    const bool same = P.model == o.model && P.integral_type == o.integral_type && P.lookup_policy == o.lookup_policy &&
                      P.nonbonded_method == o.nonbonded_method && P.cutoff == o.cutoff && P.n_radial == o.n_radial &&
                      P.n_lebedev == o.n_lebedev && P.r0 == o.r0 && P.r1 == o.r1 && P.grid_length == o.grid_length &&
                      P.buffer == o.buffer && P.cell_cap == o.cell_cap && P.orders[0] == o.orders[0] &&
                      P.orders[1] == o.orders[1] && P.tight_grid_length == o.tight_grid_length &&
                      P.block_volume == o.block_volume && P.block_force == o.block_force;
    if (!same) return fail(GBI_ERR_INVALID_ARGUMENT, "gbi_update_constants changes only the model's constants");
    ST_TRY(validate_params(P, s.N, s.custom));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    s.P = P;
    ST_TRY(upload_atoms(*ctx));  // the volume constants enter each atom's parameters and reach
    if (ctx->cell_total > 0) ST_TRY(layout(*ctx));  // the descriptors; drops captured graphs
    if (ctx->have_positions && ctx->cell_total > 0) return size_from_device(*ctx, 0);
    return GBI_OK;
}

gbi_status gbi_get_custom_values(gbi_ctx* ctx, double* values) {
    if (!ctx || !values) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->cprog) return fail(GBI_ERR_INVALID_ARGUMENT, "not a general-expression context");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "no evaluation yet");
    const size_t n = (size_t)ctx->cprog->n_values * ctx->N;
    if (n == 0) return GBI_OK;
    HIP_TRY(hipMemcpyAsync(values, ctx->c_vals.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_set_positions(gbi_ctx* ctx, const double* pos, int32_t pos_on_device) {
    if (!ctx || !pos) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(ctx->pos.p, pos, sizeof(double) * 3 * ctx->N,
                           pos_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    if (!pos_on_device) HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->have_positions = true;
    return GBI_OK;
}

gbi_status gbi_compute(gbi_ctx* ctx, uint32_t flags) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    return run(*ctx, flags);
}

gbi_status gbi_synchronize(gbi_ctx* ctx) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    bool retry = false;
    ST_TRY(check_flags(*ctx, retry));
    if (retry) return run_checked(*ctx, ctx->computed_adjoint ? GBI_ADJOINT : 0u);
    return GBI_OK;
}

gbi_status gbi_execute(gbi_ctx* ctx, const double* pos, int32_t include_forces, int32_t include_energy,
                       double* energy, double* forces) {
    if (!ctx || !pos) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (include_forces && !forces) return fail(GBI_ERR_INVALID_ARGUMENT, "null force array");
    for (int a = 0; a < 3 * ctx->N; ++a)
        if (!std::isfinite(pos[a])) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    ST_TRY(gbi_set_positions(ctx, pos, 0));
    ST_TRY(run_checked(*ctx, GBI_FORCES | GBI_ENERGY));
    if (include_energy && energy) ST_TRY(gbi_get_energy(ctx, energy));
    if (include_forces) return gbi_get_forces(ctx, forces, 1);
    return GBI_OK;
}

gbi_status gbi_execute_noalloc(int32_t n_atoms, const double* charge, const double* radius, const gbi_params* params,
                               const double* pos, int32_t include_forces, int32_t include_energy, double* energy,
                               double* forces, void* hip_stream) {
    if (!charge || !radius || !params || !pos) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (include_forces && !forces) return fail(GBI_ERR_INVALID_ARGUMENT, "null force array");
    if (params->nonbonded_method != GBI_NO_CUTOFF)
        return fail(GBI_ERR_UNSUPPORTED, "gbi_execute_noalloc evaluates NoCutoff systems (use gbi_create / gbi_execute)");
    for (long long a = 0; a < 3ll * std::max(n_atoms, 0); ++a)
        if (!std::isfinite(pos[a])) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    gbi_ctx c;
    // the buffers below are pieces of the scratch array: the guard takes them back from the
    // context before it is destroyed, so nothing is ever freed
    struct Unborrow {
        gbi_ctx& c;
        ~Unborrow() {
            auto clear = [](auto& b) {
                b.p = nullptr;
                b.n = 0;
            };
            clear(c.d_sys), clear(c.geom), clear(c.d_box), clear(c.flags_all), clear(c.work), clear(c.atom_sys);
            clear(c.bin_order), clear(c.pair_atom), clear(c.colblk), clear(c.blk_lo), clear(c.blk_hi);
            clear(c.pair_tiles), clear(c.pos), clear(c.d_charge), clear(c.d_radius), clear(c.d_reach);
            clear(c.aparam), clear(c.quad_off), clear(c.quad_pw), clear(c.quad_r2), clear(c.integrals), clear(c.born);
            clear(c.dbdi), clear(c.dedi), clear(c.atom_energy), clear(c.energy), clear(c.pair_part), clear(c.colpart);
            clear(c.force_fix), clear(c.cell_fill), clear(c.cell_count_eff), clear(c.cell_pt_count);
            clear(c.cell_ctr);
            clear(c.cell_act_count), clear(c.cell_sys), clear(c.cell_start), clear(c.frec_base), clear(c.worklist);
            clear(c.worklist_sys), clear(c.scan_tmp), clear(c.cell_atoms_tmp), clear(c.cell_atoms), clear(c.ent_pos);
            clear(c.ent_par), clear(c.prec), clear(c.frec0), clear(c.frec1), clear(c.frec2), clear(c.vval);
            clear(c.vrec), clear(c.brick_count), clear(c.brick_rec);
        }
    } unborrow{c};
    const int32_t n = n_atoms;
    ST_TRY(setup_host(c, 1, &n, charge, radius, params, false));
    HostTables t;
    ST_TRY(host_tables(c, t));
    std::vector<float4> ap;
    std::vector<double> reach;
    atom_params(c, ap, reach);
    set_max_reach(c, reach);
    std::vector<int4> tiles;
    std::vector<int> colblk;
    tile_layout(c, tiles, colblk);
    SysHost& sh = c.sys[0];
    NoAllocCaps caps;
    try {
        caps = noalloc_caps(sh, pos, radius, reach.data());
    } catch (const std::exception& e) {
        return fail(GBI_ERR_INVALID_ARGUMENT, e.what());
    }
    if (caps.cells > kMaxCells || caps.entries > INT_MAX / 2)
        return fail(GBI_ERR_GRID_TOO_LARGE, "lookup grid too large for gbi_execute_noalloc");
    sh.cell_cap = caps.cells;
    sh.tmp_cap = caps.tmp_cap;
    sh.pt_cap = caps.pt_cap;
    c.cell_total = caps.cells;
    c.entry_cap = (int)caps.entries;
    c.prec_total = caps.cells * (long long)caps.pt_cap;
    // binned by cells whatever the grid (bricks would need a second record array; the results
    // do not depend on the binning)
    c.brick_total = c.brick_max = 0;
    c.any_bricks = false;
    c.any_cells = true;

    // the buffers, carved once without a base (the size check, before the device is touched)
    // and once from the scratch array
    const size_t N = c.N, NQ = c.NQ, Q = c.quad_rows, cells = caps.cells;
    auto carve = [&](ScratchCarver& sc) {
    sc.take(c.d_sys, 1), sc.take(c.geom, 1), sc.take(c.d_box, 3), sc.take(c.flags_all, 1), sc.take(c.work, 8);
    sc.take(c.atom_sys, N), sc.take(c.bin_order, N), sc.take(c.pair_atom, N), sc.take(c.colblk, colblk.size());
    sc.take(c.blk_lo, colblk.size()), sc.take(c.blk_hi, colblk.size()), sc.take(c.pair_tiles, tiles.size());
    sc.take(c.pos, 3 * N), sc.take(c.d_charge, N), sc.take(c.d_radius, N), sc.take(c.d_reach, N), sc.take(c.aparam, N);
    sc.take(c.quad_off, 3 * Q), sc.take(c.quad_pw, 2 * Q), sc.take(c.quad_r2, Q);
    sc.take(c.integrals, 2 * N), sc.take(c.born, N), sc.take(c.dbdi, 2 * N), sc.take(c.dedi, 2 * N);
    sc.take(c.atom_energy, N), sc.take(c.energy, 1), sc.take(c.pair_part, (size_t)c.max_splits * 6 * N);
    sc.take(c.colpart, (size_t)5 * c.colpart_total), sc.take(c.force_fix, 3 * N);
    sc.take(c.cell_fill, cells), sc.take(c.cell_count_eff, cells), sc.take(c.cell_pt_count, cells);
    sc.take(c.cell_act_count, cells), sc.take(c.cell_sys, cells), sc.take(c.cell_start, cells + 1);
    sc.take(c.frec_base, cells + 1), sc.take(c.worklist, gbi::VOL_KINDS * cells);
    sc.take(c.worklist_sys, gbi::VOL_KINDS * cells), sc.take(c.scan_tmp, (cells + 4095) / 4096 + 16);
    sc.take(c.cell_atoms_tmp, cells * (size_t)caps.tmp_cap), sc.take(c.cell_atoms, caps.entries);
    sc.take(c.ent_pos, caps.entries), sc.take(c.ent_par, caps.entries), sc.take(c.prec, c.prec_total);
    sc.take(c.frec0, NQ), sc.take(c.frec1, NQ), sc.take(c.frec2, c.kinds[gbi::VOL_GBMV2] ? NQ : 1);
    sc.take(c.vval, NQ), sc.take(c.vrec, NQ);
    sc.take(c.brick_count, gbi::kBrickStride), sc.take(c.brick_rec, 1);
    sc.take(c.bbox_part, (size_t)gbi::kBboxMaxSlices * gbi::kBboxPart), sc.take(c.bbox_done, 1);
    };
    ScratchCarver dry;
    carve(dry);
    if (dry.used > gbi::kNoAllocScratchBytes)
        return fail(GBI_ERR_INVALID_ARGUMENT,
                    "system too large for gbi_execute_noalloc: it needs " + std::to_string((dry.used >> 20) + 1) +
                        " MiB of device scratch, the library holds " +
                        std::to_string(gbi::kNoAllocScratchBytes >> 20) + " MiB (use gbi_create / gbi_execute)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GBI_ERR_NO_DEVICE, "no HIP device");
    (void)hipGetDevice(&c.device);
    if (hipDeviceGetAttribute(&c.num_cus, hipDeviceAttributeMultiprocessorCount, c.device) != hipSuccess ||
        c.num_cus <= 0)
        c.num_cus = 256;
    void* base = nullptr;
    HIP_TRY(gbi::noalloc_scratch(&base));  // this device's array (looked up once per device)
    ScratchCarver sc;
    sc.base = static_cast<char*>(base);
    carve(sc);

    // one call at a time per process: every call on a GPU shares its array
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    hipStream_t st = (hipStream_t)hip_stream;
    c.stream = st;
    gbi::SysDesc desc = make_desc(c, sh, 0, 0, 0);
    desc.sk = 1;
    const double box[3] = {0.0, 0.0, 0.0};
    std::vector<int> cell_sys(cells, 0);
    auto h2d = [st](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
    };
    HIP_TRY(h2d(c.d_sys.p, &desc, sizeof(desc)));
    HIP_TRY(h2d(c.d_box.p, box, sizeof(box)));
    HIP_TRY(h2d(c.atom_sys.p, t.atom_sys.data(), sizeof(int) * N));
    HIP_TRY(h2d(c.bin_order.p, t.bin_order.data(), sizeof(int) * N));
    HIP_TRY(h2d(c.pair_atom.p, t.pair_atom.data(), sizeof(int) * N));
    HIP_TRY(h2d(c.colblk.p, colblk.data(), sizeof(int) * colblk.size()));
    HIP_TRY(h2d(c.pair_tiles.p, tiles.data(), sizeof(int4) * tiles.size()));
    HIP_TRY(h2d(c.pos.p, pos, sizeof(double) * 3 * N));
    HIP_TRY(h2d(c.d_charge.p, c.charge.data(), sizeof(double) * N));
    HIP_TRY(h2d(c.d_radius.p, c.radius.data(), sizeof(double) * N));
    HIP_TRY(h2d(c.d_reach.p, reach.data(), sizeof(double) * N));
    HIP_TRY(h2d(c.aparam.p, ap.data(), sizeof(float4) * N));
    HIP_TRY(h2d(c.quad_off.p, t.off.data(), sizeof(double) * t.off.size()));
    HIP_TRY(h2d(c.quad_pw.p, t.pw.data(), sizeof(double) * t.pw.size()));
    HIP_TRY(h2d(c.quad_r2.p, t.r2.data(), sizeof(float) * t.r2.size()));
    HIP_TRY(h2d(c.cell_sys.p, cell_sys.data(), sizeof(int) * cells));
    const gbi::KParams k = make_kparams(c, false);
    HIP_TRY(gbi::launch_evaluate(k, c.kinds, st, nullptr));
    gbi::SysGeom G;
    int all = 0;
    std::vector<long long> fx(3 * N);
    double e = 0.0;
    HIP_TRY(hipMemcpyAsync(&G, c.geom.p, sizeof(G), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&all, c.flags_all.p, sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(&e, c.energy.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipMemcpyAsync(fx.data(), c.force_fix.p, sizeof(long long) * 3 * N, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    all |= G.flags;
    if (all & gbi::FLAG_NONFINITE_POS) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    if (all & gbi::FLAG_GRID_TOO_LARGE)
        return fail(GBI_ERR_GRID_TOO_LARGE,
                    "CharmmGBMVForce: CustomGBIntegral lookup table dimension is too large, check atom positions!");
    if (all & (gbi::FLAG_CELL_OVERFLOW | gbi::FLAG_ENTRY_OVERFLOW | gbi::FLAG_POINT_OVERFLOW | gbi::FLAG_LIST_OVERFLOW))
        return fail(GBI_ERR_CAPACITY, "the lookup grid outgrew the capacities sized for it (use gbi_create / gbi_execute)");
    if (all & gbi::FLAG_NONFINITE_RESULT) return fail(GBI_ERR_NONFINITE, "non-finite Born radius, energy or force");
    if (include_energy && energy) *energy = e;
    if (include_forces)
        for (size_t a = 0; a < N; ++a)
            for (int d = 0; d < 3; ++d) forces[3 * a + d] += (double)fx[(size_t)d * N + a] / 4294967296.0;
    return GBI_OK;
}

gbi_status gbi_execute_adjoint(gbi_ctx* ctx, const double* pos, double* energy, double* forces, double* dEdR,
                               double* dEdq, double* dEdtheta) {
    if (!ctx || !pos) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    for (int a = 0; a < 3 * ctx->N; ++a)
        if (!std::isfinite(pos[a])) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    ST_TRY(gbi_set_positions(ctx, pos, 0));
    ST_TRY(run_checked(*ctx, GBI_FORCES | GBI_ENERGY | GBI_ADJOINT));
    if (energy) ST_TRY(gbi_get_energy(ctx, energy));
    if (forces) ST_TRY(gbi_get_forces(ctx, forces, 1));
    return gbi_get_adjoint(ctx, dEdR, dEdq, dEdtheta);
}

gbi_status gbi_get_energy(gbi_ctx* ctx, double* energy) {
    if (!ctx || !energy) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "nothing computed yet");
    HIP_TRY(hipMemcpyAsync(energy, ctx->energy.p, sizeof(double) * ctx->S, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_get_forces(gbi_ctx* ctx, double* forces, int32_t add) {
    if (!ctx || !forces) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "nothing computed yet");
    std::vector<long long> fx(3 * (size_t)ctx->N);
    HIP_TRY(hipMemcpyAsync(fx.data(), ctx->force_fix.p, sizeof(long long) * 3 * ctx->N, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const double scale = 1.0 / 4294967296.0;
    const int N = ctx->N;
    for (int a = 0; a < N; ++a)
        for (int d = 0; d < 3; ++d) {
            const double f = (double)fx[(size_t)d * N + a] * scale;
            if (add)
                forces[3 * a + d] += f;
            else
                forces[3 * a + d] = f;
        }
    return GBI_OK;
}

gbi_status gbi_get_adjoint(gbi_ctx* ctx, double* dEdR, double* dEdq, double* dEdtheta) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    if (!ctx->computed || !ctx->computed_adjoint)
        return fail(GBI_ERR_INVALID_ARGUMENT, "the last evaluation did not compute the adjoint");
    if (dEdR)
        HIP_TRY(hipMemcpyAsync(dEdR, ctx->dedr.p, sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    if (dEdq)
        HIP_TRY(hipMemcpyAsync(dEdq, ctx->dedq.p, sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    if (dEdtheta)
        HIP_TRY(hipMemcpyAsync(dEdtheta, ctx->theta.p, sizeof(double) * ctx->S * gbi::TH_COUNT,
                               hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_get_born_radii(gbi_ctx* ctx, double* born) {
    if (!ctx || !born) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(born, ctx->born.p, sizeof(double) * ctx->N, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_get_integrals(gbi_ctx* ctx, double* integrals) {
    if (!ctx || !integrals) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipMemcpyAsync(integrals, ctx->integrals.p, sizeof(double) * 2 * ctx->N, hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return GBI_OK;
}

gbi_status gbi_get_stats(gbi_ctx* ctx, gbi_stats* stats) {
    if (!ctx || !stats) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "nothing computed yet");
    HIP_TRY(hipMemsetAsync(ctx->stats.p, 0, sizeof(double) * 6, ctx->stream));
    gbi::KParams k = make_kparams(*ctx, false);
    HIP_TRY(gbi::launch_stats(k, ctx->stats.p, ctx->stream));
    double h[6];
    HIP_TRY(hipMemcpyAsync(h, ctx->stats.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    stats->n_points = h[0];
    stats->n_vol_pairs = h[1];
    stats->n_cells = h[2];
    stats->n_entries = h[3];
    stats->n_atoms = ctx->N;
    double nq = 0.0;  // quadrature points (NQ also counts each system's alignment padding)
    for (const SysHost& s : ctx->sys) nq += (double)s.N * s.Q;
    stats->n_quad = nq / ctx->N;
    stats->n_active = h[4];
    stats->n_frc_pairs = h[5];
    double slot = 0.0;
    for (const SysHost& s : ctx->sys) slot += (brick_edge(s) > 1 ? 4.0 : 16.0) * s.N;
    stats->slot_bytes = slot / ctx->N;
    return GBI_OK;
}

gbi_status gbi_set_graph(gbi_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->use_graph = enable != 0;
    if (!ctx->use_graph) ctx->drop_graphs();
    return GBI_OK;
}

gbi_status gbi_set_tuning(gbi_ctx* ctx, const gbi_tuning* t) {
    if (!ctx || !t) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (t->blocks_per_cu < 1 || t->blocks_per_cu > 256)
        return fail(GBI_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in 1..256");
    if (t->pair_columns < 0) return fail(GBI_ERR_INVALID_ARGUMENT, "pair_columns must be >= 0");
    if (t->tight_grid_length != 0.0 && !(t->tight_grid_length >= 0.05 && t->tight_grid_length <= 1.0))
        return fail(GBI_ERR_INVALID_ARGUMENT, "tight_grid_length must be 0 or in [0.05, 1] nm");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const bool tiles = t->pair_columns != ctx->tuning.pair_columns;
    ctx->drop_graphs();
    ctx->tuning.blocks_per_cu = t->blocks_per_cu;
    ctx->tuning.pair_columns = t->pair_columns;
    ctx->tuning.tight_grid_length = t->tight_grid_length;
    if (tiles) {
        ST_TRY(build_tiles(*ctx));
        if (ctx->cell_total > 0) ST_TRY(layout(*ctx));  // pair splits live in the descriptors
    }
    if (t->tight_grid_length > 0.0) ST_TRY(apply_tight_grid(*ctx, t->tight_grid_length));
    return GBI_OK;
}

gbi_status gbi_get_tuning(gbi_ctx* ctx, gbi_tuning* out) {
    if (!ctx || !out) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    *out = ctx->tuning;
    return GBI_OK;
}

gbi_status gbi_autotune(gbi_ctx* ctx, int32_t evals, gbi_tuning* out) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    if (!ctx->have_positions) return fail(GBI_ERR_INVALID_ARGUMENT, "positions were never set");
    if (evals < 1) evals = 3;
    bool tight = false;
    for (const SysHost& s : ctx->sys) tight |= s.kind == gbi::GRID_TIGHT;
    std::vector<gbi_tuning> cands;
    const double grids_tight[] = {kDefaultTightGrid, 0.15, 0.25, 0.3};
    for (double g : grids_tight) {
        if (!tight && g != grids_tight[0]) continue;
        for (int bpc : {16, 32, 64, 128, 256})
            for (int cols : {0, 1024, 16384}) {
                gbi_tuning t{};
                t.blocks_per_cu = bpc;
                t.pair_columns = cols;
                t.tight_grid_length = tight ? g : 0.0;
                cands.push_back(t);
            }
    }
    const bool prof = ctx->profiling;
    ctx->profiling = false;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    double best = 1e300;
    gbi_tuning best_t = ctx->tuning;
    gbi_status st = GBI_OK;
    for (const gbi_tuning& t : cands) {
        st = gbi_set_tuning(ctx, &t);
        if (st != GBI_OK) break;
        st = run_checked(*ctx, GBI_FORCES | GBI_ENERGY);  // sizes the buffers for this shape
        if (st != GBI_OK) break;
        if (hipEventRecord(e0, ctx->stream) != hipSuccess) { st = fail(GBI_ERR_HIP, "hipEventRecord"); break; }
        for (int r = 0; r < evals && st == GBI_OK; ++r) st = run(*ctx, GBI_FORCES | GBI_ENERGY);
        if (st != GBI_OK) break;
        if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
            st = fail(GBI_ERR_HIP, "hipEventSynchronize");
            break;
        }
        bool retry = false;
        st = check_flags(*ctx, retry);
        if (st != GBI_OK) break;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, e0, e1) != hipSuccess) { st = fail(GBI_ERR_HIP, "hipEventElapsedTime"); break; }
        if (!retry && ms / evals < best) {
            best = ms / evals;
            best_t = t;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->profiling = prof;
    if (st != GBI_OK) return st;
    ST_TRY(gbi_set_tuning(ctx, &best_t));
    ST_TRY(run_checked(*ctx, GBI_FORCES | GBI_ENERGY));
    ctx->tuning.ms_per_eval = best;
    if (out) *out = ctx->tuning;
    return GBI_OK;
}

gbi_status gbi_profile_enable(gbi_ctx* ctx, int32_t enable) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->profiling = enable != 0;
    ctx->prof_used = 0;
    return GBI_OK;
}

gbi_status gbi_profile_read(gbi_ctx* ctx, double* avg_ms, int32_t* launches) {
    if (!ctx || !avg_ms) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < gbi::kPasses; ++k) avg_ms[k] = 0.0;
    for (size_t l = 0; l < ctx->prof_used; ++l) {
        for (int k = 0; k < gbi::kPasses; ++k) {
            float ms = 0.f;
            HIP_TRY(hipEventElapsedTime(&ms, ctx->prof_events[l][k], ctx->prof_events[l][k + 1]));
            avg_ms[k] += ms;
        }
    }
    if (ctx->prof_used > 0)
        for (int k = 0; k < gbi::kPasses; ++k) avg_ms[k] /= (double)ctx->prof_used;
    if (launches) *launches = (int32_t)ctx->prof_used;
    return GBI_OK;
}

const char* gbi_profile_pass_name(int32_t pass) {
    static const char* names[gbi::kPasses] = {"grid", "bin", "volume", "integrals",
                                              "pairs", "chain", "force", "owner"};
    return pass >= 0 && pass < gbi::kPasses ? names[pass] : "";
}

gbi_status gbi_get_device_view(gbi_ctx* ctx, gbi_device_view* view) {
    if (!ctx || !view) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    view->positions = ctx->pos.p;
    view->forces = (int64_t*)ctx->force_fix.p;
    view->energy = ctx->energy.p;
    view->born_radii = ctx->born.p;
    view->integrals = ctx->integrals.p;
    view->stream = (void*)ctx->stream;
    return GBI_OK;
}

}  // extern "C"
