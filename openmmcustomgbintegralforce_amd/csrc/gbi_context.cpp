// Host side of the C ABI (include/gbi.h): parameter defaults, device buffers, capacity
// management and the execute sequence.  Replaces the kernel drivers
// ReferenceCalcCharmmGB{SW,MV}ForceKernel::execute (CharmmReferenceKernels.cpp:337-370,
// :426-437) and CudaCalcCharmmGBMVForceKernel::execute / CustomGBIntegral::evaluate
// (CharmmCudaKernels.cpp:845-993, cuda CustomGBIntegral.cpp:625-717).
#include "gbi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

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

constexpr double kOne4PiEps0 = 138.935456;  // OpenMM 7.x SimTKOpenMMRealType.h
constexpr double kTightMargin = 1e-6;       // nm, guards the exact GBSW lists against rounding

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t ensure(size_t count) {
        if (count <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = std::max<size_t>(count, 1);
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

}  // namespace

struct gbi_ctx {
    int N = 0;
    int Q = 0;
    gbi_params P{};
    gbi::GridKind kind = gbi::GRID_FAITHFUL;
    gbi::VolKind vol = gbi::VOL_GBSW;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    int device = 0;
    bool have_positions = false;
    bool computed = false;

    std::vector<double> charge, radius;
    int cell_cap = 0, entry_cap = 0;

    // device buffers
    DevBuf<double> pos, d_charge, d_radius, d_reach, quad_off, quad_pw, integrals, born, dbdi, dedi, pair_part,
        atom_energy, energy, stats;
    DevBuf<float4> aparam;
    DevBuf<gbi::Geom> geom;
    DevBuf<int> cell_count, cell_fill, cell_start, cell_count_eff, cell_atoms_tmp, cell_atoms, cell_pt_count,
        cell_act_count, frec_base, worklist, work, scan_tmp;
    DevBuf<float4> ent_pos, ent_par, prec, frec0, frec1, frec2, vrec;
    DevBuf<float> quad_r2;
    int pt_cap = 0;
    DevBuf<long long> force_fix;

    int pair_splits = 1;
    int perm_stride = 1;
    double gpad = 0.0, halfdiag = 0.0, max_reach = 0.0;

    // diagnostics: kPasses + 1 events per profiled evaluation
    bool profiling = false;
    std::vector<std::array<hipEvent_t, gbi::kPasses + 1>> prof_events;
    size_t prof_used = 0;

    void release_profile() {
        for (auto& a : prof_events)
            for (auto& e : a) (void)hipEventDestroy(e);
        prof_events.clear();
        prof_used = 0;
    }

    ~gbi_ctx() {
        release_profile();
        DevBuf<double>* d[] = {&pos, &d_charge, &d_radius, &d_reach, &quad_off, &quad_pw, &integrals, &born,
                               &dbdi, &dedi, &pair_part, &atom_energy, &energy, &stats};
        for (auto* b : d) b->release();
        aparam.release();
        geom.release();
        DevBuf<int>* ib[] = {&cell_count, &cell_fill, &cell_start, &cell_count_eff, &cell_atoms_tmp, &cell_atoms,
                             &cell_pt_count, &cell_act_count, &frec_base, &worklist, &work, &scan_tmp};
        for (auto* b : ib) b->release();
        DevBuf<float4>* fb[] = {&ent_pos, &ent_par, &prec, &frec0, &frec1, &frec2, &vrec};
        for (auto* b : fb) b->release();
        quad_r2.release();
        force_fix.release();
        if (own_stream) (void)hipStreamDestroy(own_stream);
    }
};

namespace {

gbi_status validate_params(const gbi_params& P, int N) {
    if (N <= 0) return fail(GBI_ERR_INVALID_ARGUMENT, "n_atoms must be positive");
    if (P.model != GBI_MODEL_GBSW && P.model != GBI_MODEL_GBMV)
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown model");
    if (P.model == GBI_MODEL_GBMV && P.integral_type == GBI_INTEGRAL_CUSTOM)
        return fail(GBI_ERR_UNSUPPORTED, "CustomIntegral (type 3) is not implemented by the reference either");
    if (P.model == GBI_MODEL_GBMV && (P.integral_type < 0 || P.integral_type > 2))
        return fail(GBI_ERR_INVALID_ARGUMENT, "unknown GB integral type");
    if (P.nonbonded_method != GBI_NO_CUTOFF)
        return fail(GBI_ERR_UNSUPPORTED, "only NoCutoff is supported on the HIP platform");
    if (P.n_radial <= 0 || P.n_radial * P.n_lebedev > (1 << gbi::kQBits) - 1)
        return fail(GBI_ERR_INVALID_ARGUMENT, "quadrature must have 1..2047 points per atom");
    if (!(P.n_lebedev == 6 || P.n_lebedev == 14 || P.n_lebedev == 26 || P.n_lebedev == 38 || P.n_lebedev == 50))
        return fail(GBI_ERR_INVALID_ARGUMENT, "n_lebedev must be 6, 14, 26, 38 or 50");
    if (!(P.r0 > 0.0) || !(P.r1 > P.r0)) return fail(GBI_ERR_INVALID_ARGUMENT, "need 0 < r0 < r1");
    if (!(P.grid_length > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "grid_length must be positive");
    for (int k = 0; k < 2; ++k)
        if (P.orders[k] < 2)  // GBSWIntegral.cpp:92-95
            return fail(GBI_ERR_INVALID_ARGUMENT, "the order of an integral must be >= 2");
    const bool o47 = (P.orders[0] == 4 && P.orders[1] == 7) || (P.orders[0] == 7 && P.orders[1] == 4);
    if (!o47) return fail(GBI_ERR_UNSUPPORTED, "the canonical Born map needs integrals of order 4 and 7");
    if ((long long)N * P.n_radial * P.n_lebedev >= (1ll << 31))
        return fail(GBI_ERR_INVALID_ARGUMENT, "N * Q must be below 2^31");
    if (N >= (1 << (32 - gbi::kQBits))) return fail(GBI_ERR_INVALID_ARGUMENT, "too many atoms");
    return GBI_OK;
}

// Per-atom constants (float for the volume functions, double reach for the tight grid).
void atom_params(const gbi_ctx& c, std::vector<float4>& ap, std::vector<double>& reach) {
    const gbi_params& P = c.P;
    ap.resize(c.N);
    reach.resize(c.N);
    for (int j = 0; j < c.N; ++j) {
        const double R = c.radius[j];
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (c.vol == gbi::VOL_GBSW) {
            const double Rp = (R + P.roffset) * P.rscale;
            const double lo = Rp - P.sw, hi = Rp + P.sw;
            v.x = (float)Rp;
            v.y = lo > 0.0 ? (float)(lo * lo) : -1.0f;
            v.z = (float)(hi * hi);
            reach[j] = hi + kTightMargin;
        } else if (c.vol == gbi::VOL_GBMV1) {
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

gbi_status upload_params(gbi_ctx& c) {
    std::vector<float4> ap;
    std::vector<double> reach;
    atom_params(c, ap, reach);
    HIP_TRY(hipMemcpyAsync(c.aparam.p, ap.data(), sizeof(float4) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_reach.p, reach.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_charge.p, c.charge.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.d_radius.p, c.radius.data(), sizeof(double) * c.N, hipMemcpyHostToDevice, c.stream));
    c.max_reach = 0.0;
    for (double r : reach) c.max_reach = std::max(c.max_reach, r);
    HIP_TRY(hipStreamSynchronize(c.stream));
    return GBI_OK;
}

// Grid sizes for the current positions (host copy) with headroom for motion.
void estimate_caps(const gbi_ctx& c, const std::vector<double>& pos, long long& cells, long long& entries) {
    if (c.kind == gbi::GRID_BRUTE) {
        cells = 1;
        entries = c.N;
        return;
    }
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = pos[d];
    for (int a = 0; a < c.N; ++a)
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(mn[d], pos[3 * a + d]);
            mx[d] = std::max(mx[d], pos[3 * a + d]);
        }
    const double g = c.P.grid_length > 0 && c.kind == gbi::GRID_FAITHFUL ? c.P.grid_length
                                                                         : c.P.tight_grid_length;
    // headroom: 0.5 nm per side plus 10 %
    const double gridpad = c.kind == gbi::GRID_FAITHFUL ? c.gpad : c.max_reach;
    cells = 1;
    for (int d = 0; d < 3; ++d) {
        const double len = (mx[d] - mn[d]) * 1.1 + 2.0 * gridpad + 1.0;
        cells *= (long long)std::ceil(len / g) + 2;
    }
    // entries: box of candidate cells per atom (an upper bound of the members)
    double per_atom;
    if (c.kind == gbi::GRID_FAITHFUL) {
        double maxR = 0.0;
        for (double r : c.radius) maxR = std::max(maxR, r);
        const double pad = maxR + c.halfdiag + c.P.buffer;
        const double side = std::floor(2.0 * pad / g) + 3.0;
        per_atom = side * side * side;
    } else {
        const double side = std::floor(2.0 * c.max_reach / g) + 3.0;
        per_atom = side * side * side;
    }
    entries = (long long)(per_atom * c.N * 1.05) + 1024;
}

gbi_status ensure_grid_buffers(gbi_ctx& c, long long cells, long long entries) {
    if (cells > (1ll << 30) || entries > (1ll << 31) - 1)
        return fail(GBI_ERR_GRID_TOO_LARGE, "lookup grid too large for the device buffers");
    if (cells > 16ll * 1024 * 1024 * 4) return fail(GBI_ERR_GRID_TOO_LARGE, "more than 64M lookup cells");
    if (cells > c.cell_cap) {
        c.cell_cap = (int)cells;
        HIP_TRY(c.cell_count.ensure(cells));
        HIP_TRY(c.cell_fill.ensure(cells));
        HIP_TRY(c.cell_start.ensure(cells + 1));
        HIP_TRY(c.cell_count_eff.ensure(cells));
        HIP_TRY(c.cell_pt_count.ensure(cells));
        HIP_TRY(c.cell_act_count.ensure(cells));
        HIP_TRY(c.frec_base.ensure(cells + 1));
        HIP_TRY(c.worklist.ensure(cells));
        const long long tiles = (cells + 4095) / 4096;
        HIP_TRY(c.scan_tmp.ensure(std::max<long long>(tiles, 1) + 16));
    }
    HIP_TRY(c.prec.ensure((size_t)c.cell_cap * c.pt_cap));
    if (entries > c.entry_cap) {
        c.entry_cap = (int)entries;
        HIP_TRY(c.cell_atoms_tmp.ensure(entries));
        HIP_TRY(c.cell_atoms.ensure(entries));
        HIP_TRY(c.ent_pos.ensure(entries));
        HIP_TRY(c.ent_par.ensure(entries));
    }
    return GBI_OK;
}

gbi::KParams make_kparams(gbi_ctx& c) {
    gbi::KParams k{};
    const gbi_params& P = c.P;
    k.N = c.N;
    k.Q = c.Q;
    k.NQ = c.N * c.Q;
    k.cell_cap = c.cell_cap;
    k.entry_cap = c.entry_cap;
    k.grid_kind = c.kind;
    k.vol_kind = c.vol;
    k.cap = P.cell_cap > 0 ? P.cell_cap : INT32_MAX;
    k.n_lebedev = P.n_lebedev;
    k.g = c.kind == gbi::GRID_FAITHFUL ? P.grid_length : P.tight_grid_length;
    k.inv_g = 1.0 / k.g;
    k.gpad = c.kind == gbi::GRID_FAITHFUL ? c.gpad : c.max_reach;
    k.halfdiag = c.halfdiag;
    k.buffer = P.buffer;
    k.tight_margin = kTightMargin;
    k.sw_d = P.sw;
    k.sw = (float)P.sw;
    k.k1 = (float)(3.0 / (4.0 * P.sw));
    k.k3 = (float)(1.0 / (4.0 * P.sw * P.sw * P.sw));
    k.beta = (float)P.beta;
    k.lambda = (float)P.lambda;
    k.S0 = (float)P.S0;
    k.vguard = (float)P.vguard;
    k.fclamp = 1.0e6f;
    k.ssat = 6.0f;
    k.v_empty = c.vol == gbi::VOL_GBSW ? 1.0f
                : c.vol == gbi::VOL_GBMV2 ? 0.0f
                                          : (float)(1.0 / (1.0 + std::exp(P.beta * (0.0 - P.lambda))));
    k.sign_chain = c.vol == gbi::VOL_GBSW ? 1.0f : -1.0f;
    k.alpha0 = P.alpha0;
    k.alpha1 = P.alpha1;
    k.eta = P.eta;
    k.slope = P.slope;
    k.shift = P.shift;
    k.prefactor = P.prefactor;
    k.orders[0] = P.orders[0];
    k.orders[1] = P.orders[1];
    k.i4 = P.orders[0] == 4 ? 0 : 1;
    k.i7 = 1 - k.i4;
    k.pair_splits = c.pair_splits;
    k.pos = c.pos.p;
    k.charge = c.d_charge.p;
    k.radius = c.d_radius.p;
    k.reach = c.d_reach.p;
    k.aparam = c.aparam.p;
    k.quad_off = c.quad_off.p;
    k.quad_pw = c.quad_pw.p;
    k.geom = c.geom.p;
    k.cell_count = c.cell_count.p;
    k.cell_fill = c.cell_fill.p;
    k.cell_start = c.cell_start.p;
    k.cell_count_eff = c.cell_count_eff.p;
    k.cell_atoms_tmp = c.cell_atoms_tmp.p;
    k.cell_atoms = c.cell_atoms.p;
    k.cell_pt_count = c.cell_pt_count.p;
    k.ent_pos = c.ent_pos.p;
    k.ent_par = c.ent_par.p;
    k.cell_act_count = c.cell_act_count.p;
    k.frec_base = c.frec_base.p;
    k.pt_cap = c.pt_cap;
    k.worklist = c.worklist.p;
    k.prec = c.prec.p;
    k.frec0 = c.frec0.p;
    k.frec1 = c.frec1.p;
    k.frec2 = c.frec2.p;
    k.vrec = c.vrec.p;
    k.quad_r2 = c.quad_r2.p;
    k.perm_stride = c.perm_stride;
    k.integrals = c.integrals.p;
    k.born = c.born.p;
    k.dbdi = c.dbdi.p;
    k.dedi = c.dedi.p;
    k.pair_part = c.pair_part.p;
    k.atom_energy = c.atom_energy.p;
    k.force_fix = c.force_fix.p;
    k.energy = c.energy.p;
    k.work = c.work.p;
    k.scan_tmp = c.scan_tmp.p;
    return k;
}

gbi_status read_geom(gbi_ctx& c, gbi::Geom& G) {
    HIP_TRY(hipMemcpyAsync(&G, c.geom.p, sizeof(G), hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    return GBI_OK;
}

// Size the grid buffers from the device's own bounding box (first call or after overflow).
gbi_status size_from_device(gbi_ctx& c) {
    std::vector<double> host(3 * (size_t)c.N);
    HIP_TRY(hipMemcpyAsync(host.data(), c.pos.p, sizeof(double) * 3 * c.N, hipMemcpyDeviceToHost, c.stream));
    HIP_TRY(hipStreamSynchronize(c.stream));
    for (double v : host)
        if (!std::isfinite(v)) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    long long cells, entries;
    estimate_caps(c, host, cells, entries);
    return ensure_grid_buffers(c, std::max<long long>(cells, c.cell_cap), std::max<long long>(entries, c.entry_cap));
}

gbi_status check_flags(gbi_ctx& c, bool& retry) {
    retry = false;
    gbi::Geom G;
    gbi_status st = read_geom(c, G);
    if (st != GBI_OK) return st;
    if (G.flags & gbi::FLAG_NONFINITE_POS) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    if (G.flags & gbi::FLAG_GRID_TOO_LARGE)
        return fail(GBI_ERR_GRID_TOO_LARGE,
                    "CharmmGBMVForce: CustomGBIntegral lookup table dimension is too large, check atom positions!");
    if (G.flags & (gbi::FLAG_CELL_OVERFLOW | gbi::FLAG_ENTRY_OVERFLOW | gbi::FLAG_POINT_OVERFLOW)) {
        long long cells = (long long)c.cell_cap, entries = c.entry_cap;
        if (G.flags & gbi::FLAG_POINT_OVERFLOW) c.pt_cap = std::min(2 * c.pt_cap, c.N * c.Q);
        if (G.flags & gbi::FLAG_CELL_OVERFLOW) cells = (long long)(c.cell_cap * 2ll);
        if (G.flags & gbi::FLAG_ENTRY_OVERFLOW) entries = std::max<long long>(2ll * c.entry_cap, (long long)G.total_entries * 5 / 4);
        st = size_from_device(c);
        if (st != GBI_OK) return st;
        st = ensure_grid_buffers(c, cells, entries);
        if (st != GBI_OK) return st;
        retry = true;
        return GBI_OK;
    }
    if (G.flags & gbi::FLAG_NONFINITE_RESULT) return fail(GBI_ERR_NONFINITE, "non-finite Born radius or energy");
    return GBI_OK;
}

gbi_status run(gbi_ctx& c) {
    if (!c.have_positions) return fail(GBI_ERR_INVALID_ARGUMENT, "positions were never set");
    if (c.cell_cap == 0) {
        gbi_status st = size_from_device(c);
        if (st != GBI_OK) return st;
    }
    gbi::KParams k = make_kparams(c);
    const hipEvent_t* marks = nullptr;
    if (c.profiling && c.prof_used < 4096) {
        if (c.prof_used == c.prof_events.size()) {
            std::array<hipEvent_t, gbi::kPasses + 1> a;
            for (auto& e : a) HIP_TRY(hipEventCreate(&e));
            c.prof_events.push_back(a);
        }
        marks = c.prof_events[c.prof_used++].data();
    }
    HIP_TRY(gbi::launch_evaluate(k, GBI_FORCES | GBI_ENERGY, c.stream, marks));
    c.computed = true;
    return GBI_OK;
}

gbi_status run_checked(gbi_ctx& c) {
    for (int attempt = 0; attempt < 4; ++attempt) {
        gbi_status st = run(c);
        if (st != GBI_OK) return st;
        bool retry = false;
        st = check_flags(c, retry);
        if (st != GBI_OK) return st;
        if (!retry) return GBI_OK;
    }
    return fail(GBI_ERR_CAPACITY, "lookup grid kept outgrowing its buffers");
}

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
    P.tight_grid_length = 0.15;
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
    if (!out || !charge || !radius || !params) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    gbi_status st = validate_params(*params, n_atoms);
    if (st != GBI_OK) return st;
    for (int j = 0; j < n_atoms; ++j) {
        if (!std::isfinite(charge[j]) || !std::isfinite(radius[j]))
            return fail(GBI_ERR_NONFINITE, "non-finite charge or radius");
        if (!(radius[j] > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "radii must be positive");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(GBI_ERR_NO_DEVICE, "no HIP device");
    auto* c = new gbi_ctx();
    c->N = n_atoms;
    c->P = *params;
    if (c->P.model == GBI_MODEL_GBSW) c->P.integral_type = GBI_INTEGRAL_GBSW;
    if (c->P.tight_grid_length <= 0.0) c->P.tight_grid_length = 0.15;
    c->Q = c->P.n_radial * c->P.n_lebedev;
    c->charge.assign(charge, charge + n_atoms);
    c->radius.assign(radius, radius + n_atoms);
    c->vol = (gbi::VolKind)c->P.integral_type;
    if (c->P.lookup_policy == GBI_LOOKUP_BRUTE)
        c->kind = gbi::GRID_BRUTE;
    else if (c->P.model == GBI_MODEL_GBSW)
        c->kind = gbi::GRID_TIGHT;  // the standalone lists are complete, so exact lists give the same V
    else
        c->kind = gbi::GRID_FAITHFUL;
    c->halfdiag = std::sqrt(3.0) / 2.0 * c->P.grid_length;
    {
        const double gc = c->kind == gbi::GRID_FAITHFUL ? c->P.grid_length : c->P.tight_grid_length;
        const double est = 1.6 * 100.0 * c->Q * gc * gc * gc + 64.0;
        c->pt_cap = c->kind == gbi::GRID_BRUTE ? n_atoms * c->Q : std::min((int)est, n_atoms * c->Q);
        c->pt_cap = std::max(c->pt_cap, 1);
    }
    c->gpad = 0.03 + c->halfdiag + c->P.buffer;  // CustomGBIntegral.cpp:174-175
    (void)hipGetDevice(&c->device);
    auto cleanup = [&](gbi_status s) {
        delete c;
        return s;
    };
#define CTX_TRY(expr)                                                                                \
    do {                                                                                             \
        hipError_t e_ = (expr);                                                                      \
        if (e_ != hipSuccess) {                                                                      \
            fail(e_ == hipErrorOutOfMemory ? GBI_ERR_OUT_OF_MEMORY : GBI_ERR_HIP,                    \
                 std::string(#expr) + ": " + hipGetErrorString(e_));                                 \
            return cleanup(e_ == hipErrorOutOfMemory ? GBI_ERR_OUT_OF_MEMORY : GBI_ERR_HIP);          \
        }                                                                                            \
    } while (0)
    CTX_TRY(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
    c->stream = c->own_stream;
    const size_t N = n_atoms, Q = c->Q, NQ = N * Q;
    CTX_TRY(c->pos.ensure(3 * N));
    CTX_TRY(c->d_charge.ensure(N));
    CTX_TRY(c->d_radius.ensure(N));
    CTX_TRY(c->d_reach.ensure(N));
    CTX_TRY(c->aparam.ensure(N));
    CTX_TRY(c->quad_off.ensure(3 * Q));
    CTX_TRY(c->quad_pw.ensure(2 * Q));
    CTX_TRY(c->integrals.ensure(2 * N));
    CTX_TRY(c->born.ensure(N));
    CTX_TRY(c->dbdi.ensure(2 * N));
    CTX_TRY(c->dedi.ensure(2 * N));
    CTX_TRY(c->atom_energy.ensure(N));
    CTX_TRY(c->energy.ensure(1));
    CTX_TRY(c->stats.ensure(4));
    CTX_TRY(c->geom.ensure(1));
    CTX_TRY(c->work.ensure(4));
    CTX_TRY(c->frec0.ensure(NQ));
    CTX_TRY(c->frec1.ensure(NQ));
    CTX_TRY(c->frec2.ensure(c->vol == gbi::VOL_GBMV2 ? NQ : 1));
    CTX_TRY(c->vrec.ensure(NQ));
    CTX_TRY(c->quad_r2.ensure(Q));
    CTX_TRY(c->force_fix.ensure(3 * N));
    // pair pass: enough workgroups to fill the device, at least 64 columns per split
    const int rows = (int)((N + 63) / 64);
    int splits = (2048 + rows - 1) / rows;
    splits = std::max(1, std::min(splits, std::min(64, (int)(N / 64))));
    c->pair_splits = std::max(1, splits);
    CTX_TRY(c->pair_part.ensure((size_t)c->pair_splits * 5 * N));
    // quadrature tables: offsets and w_q / r_q^(order-2) per integral
    std::vector<std::array<double, 5>> table;
    try {
        table = gbi::quadrature_table(c->P.n_radial, c->P.r0, c->P.r1, c->P.n_lebedev);
    } catch (const std::exception& e) {
        fail(GBI_ERR_INVALID_ARGUMENT, e.what());
        return cleanup(GBI_ERR_INVALID_ARGUMENT);
    }
    std::vector<double> off(3 * Q), pw(2 * Q);
    for (size_t q = 0; q < Q; ++q) {
        for (int d = 0; d < 3; ++d) off[3 * q + d] = table[q][d];
        for (int k = 0; k < 2; ++k) pw[k * Q + q] = table[q][4] / std::pow(table[q][3], c->P.orders[k] - 2);
    }
    CTX_TRY(hipMemcpyAsync(c->quad_off.p, off.data(), sizeof(double) * 3 * Q, hipMemcpyHostToDevice, c->stream));
    CTX_TRY(hipMemcpyAsync(c->quad_pw.p, pw.data(), sizeof(double) * 2 * Q, hipMemcpyHostToDevice, c->stream));
    std::vector<float> r2(Q);
    for (size_t q = 0; q < Q; ++q) r2[q] = (float)(table[q][3] * table[q][3]);
    CTX_TRY(hipMemcpyAsync(c->quad_r2.p, r2.data(), sizeof(float) * Q, hipMemcpyHostToDevice, c->stream));
    // binning visits atoms with a stride coprime to N so concurrent workgroups hit distant cells
    c->perm_stride = 1;
    for (int cand : {7919, 7927, 7933, 7937, 7949})
        if (n_atoms % cand != 0) {
            c->perm_stride = cand % n_atoms == 0 ? 1 : cand;
            break;
        }
    CTX_TRY(hipStreamSynchronize(c->stream));
#undef CTX_TRY
    st = upload_params(*c);
    if (st != GBI_OK) return cleanup(st);
    *out = c;
    return GBI_OK;
}

void gbi_destroy(gbi_ctx* ctx) {
    if (!ctx) return;
    (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
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
    for (int j = 0; j < n_atoms; ++j) {
        if (!std::isfinite(charge[j]) || !std::isfinite(radius[j]))
            return fail(GBI_ERR_NONFINITE, "non-finite charge or radius");
        if (!(radius[j] > 0.0)) return fail(GBI_ERR_INVALID_ARGUMENT, "radii must be positive");
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->charge.assign(charge, charge + n_atoms);
    ctx->radius.assign(radius, radius + n_atoms);
    gbi_status st = upload_params(*ctx);
    if (st != GBI_OK) return st;
    // radii change the lookup reach: re-estimate the grid on the next call
    if (ctx->have_positions) return size_from_device(*ctx);
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
    (void)flags;
    return run(*ctx);
}

gbi_status gbi_synchronize(gbi_ctx* ctx) {
    if (!ctx) return fail(GBI_ERR_INVALID_ARGUMENT, "null context");
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    bool retry = false;
    gbi_status st = check_flags(*ctx, retry);
    if (st != GBI_OK) return st;
    if (retry) return run_checked(*ctx);
    return GBI_OK;
}

gbi_status gbi_execute(gbi_ctx* ctx, const double* pos, int32_t include_forces, int32_t include_energy,
                       double* energy, double* forces) {
    if (!ctx || !pos) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (include_forces && !forces) return fail(GBI_ERR_INVALID_ARGUMENT, "null force array");
    for (int a = 0; a < 3 * ctx->N; ++a)
        if (!std::isfinite(pos[a])) return fail(GBI_ERR_NONFINITE, "non-finite atom position");
    gbi_status st = gbi_set_positions(ctx, pos, 0);
    if (st != GBI_OK) return st;
    st = run_checked(*ctx);
    if (st != GBI_OK) return st;
    if (include_energy && energy) {
        st = gbi_get_energy(ctx, energy);
        if (st != GBI_OK) return st;
    }
    if (include_forces) return gbi_get_forces(ctx, forces, 1);
    return GBI_OK;
}

gbi_status gbi_get_energy(gbi_ctx* ctx, double* energy) {
    if (!ctx || !energy) return fail(GBI_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->computed) return fail(GBI_ERR_INVALID_ARGUMENT, "nothing computed yet");
    HIP_TRY(hipMemcpyAsync(energy, ctx->energy.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
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
    HIP_TRY(hipMemsetAsync(ctx->stats.p, 0, sizeof(double) * 4, ctx->stream));
    gbi::KParams k = make_kparams(*ctx);
    HIP_TRY(gbi::launch_stats(k, ctx->stats.p, ctx->stream));
    double h[4];
    HIP_TRY(hipMemcpyAsync(h, ctx->stats.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    stats->n_points = h[0];
    stats->n_vol_pairs = h[1];
    stats->n_cells = h[2];
    stats->n_entries = h[3];
    stats->n_atoms = ctx->N;
    stats->n_quad = ctx->Q;
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
