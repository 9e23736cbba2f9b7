// Internal declarations shared by the gfx950 kernels (gbi_kernels.hip) and the host
// orchestration (gbi_context.cpp).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gbi {

// How the neighbour lists of quadrature points are built.
enum GridKind : int {
    GRID_FAITHFUL = 0,  // CustomGBIntegral lookup grid: corner-anchored membership, cap in index order
    GRID_TIGHT = 1,     // exact GBSW lists: atoms whose switching sphere reaches the cell box
    GRID_BRUTE = 2      // one cell holding every atom
};

// Volume function of the integral.
enum VolKind : int { VOL_GBSW = 0, VOL_GBMV1 = 1, VOL_GBMV2 = 2 };

// Device-side flags (bit set in Geom::flags).
enum : int {
    FLAG_CELL_OVERFLOW = 1,   // more cells than allocated
    FLAG_ENTRY_OVERFLOW = 2,  // more cell-list entries than allocated
    FLAG_GRID_TOO_LARGE = 4,  // an axis longer than 1000 nm (CustomGBIntegral.cpp:182-183)
    FLAG_NONFINITE_POS = 8,
    FLAG_NONFINITE_RESULT = 16,
    FLAG_POINT_OVERFLOW = 32  // more binned points than allocated (never expected)
};

// Geometry of the lookup grid, written on the device every evaluation.
struct Geom {
    double origin[3];  // grid minimum corner (brute: molecule centre)
    double max[3];     // last grid point (faithful) / upper edge (tight)
    int n[3];
    int ncell;
    int flags;
    int total_entries;
    int total_points;
    int pad_;
};

constexpr int kMaxAtomChunk = 256;   // atoms staged in LDS per chunk (volume pass)
constexpr int kQBits = 11;           // packed point = (atom << 11) | q, Q <= 2047
constexpr float kQuantum = 16777216.0f;  // 2^24: chain-rule partials are summed exactly on this grid

// Per-atom float parameters of the volume function (host-precomputed).
//   GBSW : x = R', y = (R' - w)^2, z = (R' + w)^2, w = 0
//   GBMV1: x = gamma_j = gamma0 ln(lambda) / R^4
//   GBMV2: x = C_j, y = C_j - 100 R^2, z = 400 C_j^2
struct KParams {
    // sizes
    int N, Q, NQ;
    int cell_cap, entry_cap;
    int grid_kind, vol_kind;
    int cap;               // atoms kept per cell
    int n_lebedev;
    // grid constants (double, computed on the host with the oracle's expressions)
    double g;              // cell length
    double inv_g;          // tight grid: 1/g
    double gpad;           // faithful: 0.03 + sqrt(3)/2 g + buffer; tight: max reach
    double halfdiag;       // sqrt(3.0)/2.0*g
    double buffer;
    double tight_margin;
    double sw_d;           // GBSW switching half-width (double)
    // volume constants (float)
    float sw, k1, k3;      // H = 1/2 + x (k1 - k3 x^2), k1 = 3/(4w), k3 = 1/(4w^3)
    float beta, lambda, S0, vguard, fclamp, ssat;
    float v_empty;
    float sign_chain;      // +1 for GBSW (I accumulates 1 - V), -1 for GBMV
    // Born map / energy (double)
    double alpha0, alpha1, eta, slope, shift, prefactor;
    int orders[2];
    int i4, i7;            // which integral is order 4 / 7
    int pair_splits;       // column splits of the pair pass
    // inputs
    const double* pos;     // N x 3
    const double* charge;  // N
    const double* radius;  // N
    const double* reach;   // N: tight grid reach R' + w + margin (double)
    const float4* aparam;  // N
    const double* quad_off;  // 3Q
    const double* quad_pw;   // 2Q: w_q / r_q^(order_k - 2)
    // grid
    Geom* geom;
    int* cell_count;       // cell_cap
    int* cell_fill;        // cell_cap
    int* cell_start;       // cell_cap + 1
    int* cell_count_eff;   // cell_cap
    int* cell_atoms_tmp;   // entry_cap
    int* cell_atoms;       // entry_cap
    float4* ent_pos;       // entry_cap: atom position relative to the cell corner, .w = atom index bits
    float4* ent_par;       // entry_cap: per-atom volume parameters (aparam)
    // points
    int pt_cap;            // point slots per cell
    int* cell_pt_count;    // cell_cap
    int* cell_act_count;   // cell_cap: points with a nonzero chain-rule factor
    int* frec_base;        // cell_cap + 1: first record of the cell's active points (scan of counts)
    int* worklist;         // cell_cap: cells holding points
    float4* prec;          // cell_cap x pt_cap: point relative to the cell corner, .w = packed index bits
    float4* frec0;         // NQ (active points, grouped by cell): copy of prec
    float4* frec1;         // NQ: GBSW (V, 0, 0, 0) / GBMV1 (dV/dS, V, 0, 0) / GBMV2 (c1, c2, c3, vx)
    float4* frec2;         // NQ: GBMV2 (vy, vz, V, 0)
    float4* vrec;          // NQ by (atom, q): (V, owner vector) -- GBSW dV/dp, GBMV owner force
    const float* quad_r2;  // Q: |offset_q|^2
    int perm_stride;       // atom visiting order of the binning pass
    // per atom
    double* integrals;     // 2 x N
    double* born;          // N
    double* dbdi;          // 2 x N
    double* dedi;          // 2 x N
    double* pair_part;     // pair_splits x 5 x N
    double* atom_energy;   // N
    long long* force_fix;  // 3 x N, scale 2^32
    double* energy;        // 1
    int* work;             // work counters [4]
    int* scan_tmp;         // scan block sums
};

// Launchers (gbi_kernels.hip).  All enqueue on `stream`.
// marks: nullptr or kPasses + 1 events recorded before each pass and after the last.
constexpr int kPasses = 8;
hipError_t launch_evaluate(const KParams& p, unsigned flags, hipStream_t stream, const hipEvent_t* marks);
hipError_t launch_bbox_only(const KParams& p, hipStream_t stream);
hipError_t launch_stats(const KParams& p, double* out4, hipStream_t stream);

}  // namespace gbi
