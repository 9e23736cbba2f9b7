// Internal declarations shared by the gfx950 kernels (gbi_kernels.hip) and the host
// orchestration (gbi_context.cpp).  Not part of the public ABI.
//
// A context evaluates a batch of S >= 1 independent systems with one launch per pass:
// atoms, quadrature points, lookup cells and cell-list entries of all systems live in
// shared arrays, each system owning a static range of atoms, points, cells and slots.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace gbi {

// How the neighbour lists of quadrature points are built.
enum GridKind : int {
    GRID_FAITHFUL = 0,  // CustomGBIntegral lookup grid: corner-anchored membership, cap in index order
    GRID_TIGHT = 1,     // exact GBSW lists: atoms whose switching sphere reaches the cell box
    GRID_BRUTE = 2      // one cell holding every atom of the system
};

// Volume function of the integral.
enum VolKind : int { VOL_GBSW = 0, VOL_GBMV1 = 1, VOL_GBMV2 = 2, VOL_KINDS = 3 };

// Device-side flags (SysGeom::flags of the system, OR-ed into *flags_all).
enum : int {
    FLAG_CELL_OVERFLOW = 1,    // more cells than the system's range
    FLAG_ENTRY_OVERFLOW = 2,   // more cell-list entries than allocated
    FLAG_GRID_TOO_LARGE = 4,   // an axis longer than 1000 nm (CustomGBIntegral.cpp:182-183)
    FLAG_NONFINITE_POS = 8,
    FLAG_NONFINITE_RESULT = 16,
    FLAG_POINT_OVERFLOW = 32,  // more points in a cell than its slots
    FLAG_LIST_OVERFLOW = 64    // more raw members in a cell than its list slots
};
constexpr int kFatalFlags =
    FLAG_CELL_OVERFLOW | FLAG_ENTRY_OVERFLOW | FLAG_GRID_TOO_LARGE | FLAG_NONFINITE_POS | FLAG_POINT_OVERFLOW |
    FLAG_LIST_OVERFLOW;

constexpr int kMaxAtomChunk = 256;       // atoms staged in LDS per chunk (volume pass)
constexpr int kQBits = 11;               // packed point = (atom << 11) | q, Q <= 2047
constexpr int kMaxAtoms = 1 << (32 - kQBits);
constexpr float kQuantum = 16777216.0f;  // 2^24: chain-rule partials are summed exactly on this grid
constexpr int kBrickStride = 32;         // ints between brick counters (one 128-byte line each)
constexpr int kMaxBrickEdge = 4;         // cells per brick edge (<= 64 cells per brick)
// Bounding box: a system's atoms in up to kBboxMaxSlices slices of about kBboxSlice atoms,
// one workgroup each (the count depends on the system's size only).
constexpr int kBboxSlice = 1024;
constexpr int kBboxMaxSlices = 64;
constexpr int kBboxPart = 10;            // per slice: min xyz, max xyz, sum xyz, any non-finite
__host__ __device__ inline int bbox_slices(int n_atoms) {
    const int b = (n_atoms + kBboxSlice - 1) / kBboxSlice;
    return b < 1 ? 1 : (b > kBboxMaxSlices ? kBboxMaxSlices : b);
}

// Derivatives of the energy with respect to the constants of a system (adjoint).
enum Theta : int {
    TH_PREFACTOR = 0,
    TH_ALPHA0,
    TH_ALPHA1,
    TH_ETA,
    TH_SLOPE,
    TH_SHIFT,
    TH_SW,       // GBSW switching half-width w
    TH_RSCALE,   // GBSW R' = rscale (R + roffset)
    TH_ROFFSET,
    TH_GAMMA0,   // GBMV-I
    TH_BETA,     // GBMV-I/II
    TH_LAMBDA,   // GBMV-I/II
    TH_P1,       // GBMV-II
    TH_P2,
    TH_S0,
    TH_COUNT
};
// Force-pass contributions to dE/dtheta, accumulated per system in fixed point 2^32 as
// sums over points p and list atoms j of s fac_p dV_p/d(.) (s = +1 GBSW, -1 GBMV):
enum VolAcc : int {
    VA_W = 0,   // dV/dw = V sum_j (dH_j/dw) / H_j                  (GBSW)
    VA_A,       // dV/dP2 = sum_j dV/dF_j (dF_j/dC_j + dF_j/db_j)    (GBMV-II, b_j = C_j - 100 R_j^2)
    VA_AR,      // dV/dP1 = same terms times 10 R_j                  (GBMV-II)
    VA_GAMMA,   // sum_j dV/dgamma_j gamma_j                         (GBMV-I)
    VA_COUNT
};
// The adjoint force pass adds each cell's sums to copy (worklist index mod kVaccCopies) of its
// system's accumulators, each copy on a 128-byte line of its own, and k_theta_reduce adds the
// copies: one set per system serialised every cell's atomics on a few addresses (0.21 against
// 0.16 ms for the pass at 8 x 2,000 GBSW atoms).  Fixed-point integers: any order, one sum.
constexpr int kVaccCopies = 64;
constexpr int kVaccStride = 16;
static_assert(VA_COUNT <= kVaccStride, "one 128-byte line per copy");

// Static description of one system (host-built, uploaded once per layout).
struct SysDesc {
    int grid_kind, vol_kind;
    int atom_begin, atom_end;  // global atom range
    int Q;                     // quadrature points per atom
    int quad_base;             // first row of this system's quadrature tables
    int pt_base;               // pid = pt_base + (i - atom_begin) * Q + q
    int cell_base, cell_cap;   // global cells [cell_base, cell_base + cell_cap)
    int pt_cap;                // point slots per cell
    long long prec_base;       // slot s of local cell c: prec_base + c * pt_cap + s
    int sk;                    // bricks of sk x sk x sk cells reserve point slots together (binning; 1: cells)
    int brick_base;            // the system's first brick in KParams::brick_count
    long long tmp_base;        // raw member k of local cell c: cell_atoms_tmp[tmp_base + c * tmp_cap + k]
    int tmp_cap;               // raw member slots per cell (the system's size for the brute cell)
    int cap;                   // atoms kept per cell
    int pair_splits;           // pair tiles of row block 0 (the most of any block)
    int pair_k;                // column blocks per pair tile
    int nblk;                  // 64-atom blocks
    long long colpart_base;    // column partials (row block I < block J of atom j): colpart_base + 64 (J(J-1)/2 + I) + j % 64
    int i4, i7;                // which integral is order 4 / 7
    // nonbonded method
    int periodic;              // rectangular periodic box (KParams::box), lookup grid tiling it
    int cut_mode;              // 0: no cutoff; 1: pairs with r <= c (standalone GBSW); 2: r < c (CharmmGBMVForce)
    int chain_trunc;           // standalone GBSW with a cutoff: dB_i/dx_j reaches only atoms j within c of i
    int blk_base;              // index of the system's first 64-atom block in KParams::colblk
    int custom;                // general expressions: no Born map, generated pair and chain kernels
    double cut2;               // c^2
    float cut2_lo, cut2_hi;    // FP32 band around c^2 inside which the FP64 distance decides
    float trunc_in, trunc_out; // chain_trunc: points nearer their owner than trunc_in keep every shell
                               // atom, points farther than trunc_out none (FLT_MAX when periodic)
    // grid constants (double, computed on the host with the oracle's expressions)
    double g, inv_g, gpad, halfdiag, buffer;
    // volume constants
    float k1, k3;              // GBSW: H = 1/2 + x (k1 - k3 x^2), k1 = 3/(4w), k3 = 1/(4w^3)
    float sw, inv_sw, rscale, roffset;
    float gamma0, beta, lambda, S0, vguard, fclamp, P1;
    float v_empty;             // V of a point with an empty list
    int acc64;                 // GBMV: the sigmoid's S summed in FP64 (steep sigmoids at large S, |beta lambda| > 16)
    double beta64, lambda64, S064;  // the constants of that FP64 evaluation
    float sign_chain;          // +1 for GBSW (I accumulates 1 - V), -1 for GBMV
    // Born map / energy
    double alpha0, alpha1, eta, slope, shift, prefactor;
};

// Launch arguments of the generated kernels of a system with general expressions
// (gbi_custom.h); host-side, read by launch_evaluate.
struct CustomLaunch {
    hipFunction_t values = nullptr, pairs = nullptr, chain = nullptr;
    int N = 0, nb = 0, method = 0;  // nb: 64-atom blocks
    int pair_threads = 64;          // block size of the pair kernel
    int chain_threads = 64;         // block size of the chain kernel (one per 64-atom block)
    double cut2 = 0.0;
    const double* box = nullptr;  // device: the system's 3 box edges
    const double* prm = nullptr;
    const double* glob = nullptr;
    double* vals = nullptr;
    double* dvdi = nullptr;
    double* esp = nullptr;
    double* despdi = nullptr;
    double* part = nullptr;
    double* dvdth = nullptr;   // dV/dtheta per value, derivative and atom
    double* despdth = nullptr; // single-particle dE/dtheta per atom
    double* pdth = nullptr;    // each atom's share of dE/dtheta
    double* dvdx = nullptr;    // dV/d(x, y, z) per value, axis and atom
    double* despdx = nullptr;  // d(single-particle energy)/d(x, y, z) per axis and atom
    const int* ex_ptr = nullptr;
    const int* ex_idx = nullptr;
};

// Grid geometry of one system, written on the device every evaluation.
struct SysGeom {
    double origin[3];  // grid minimum corner (brute: molecule centre)
    double max[3];     // last grid point (faithful) / upper edge (tight)
    double step[3];    // periodic: cell edges L / n
    double L[3];       // periodic: box edges
    double halfdiag;   // periodic faithful grid: sqrt3/2 max(step)
    int n[3];
    int ncell;
    int flags;
    int nb[3];         // bricks per axis: ceil(n / sk)
};

struct KParams {
    int S, N, NQ;            // systems, atoms, quadrature points (all systems)
    int cell_total;          // sum of the systems' cell ranges
    int entry_cap;
    int n_tiles;             // pair tiles
    int max_splits;          // pair_part planes (row pieces of block 0, the most of any block)
    int n_colblocks;         // 64-atom blocks of all systems (k_atom_chain workgroups)
    long long colpart_total; // column partial entries per plane
    int adjoint;             // also compute dE/dR, dE/dq, dE/dtheta
    const int* bin_order;    // N: atom visiting order of the binning pass (systems in turn, each
                             // with a stride coprime to its size so concurrent workgroups hit distant cells)
    int persist_blocks;      // workgroups of the per-cell passes
    const SysDesc* sys;
    SysGeom* geom;
    const double* box;       // S x 3: edges of each periodic system's box
    int any_cut;             // some system has a cutoff: pair tiles test distances
    int any_trunc;           // some system truncates the chain rule at the cutoff
    float4* blk_lo;          // n_colblocks: bounding box of each 64-atom block (cutoff systems)
    float4* blk_hi;
    const CustomLaunch* custom;  // host pointer: the generated kernels replace the pair and chain passes
    int any_excl;            // some system has excluded pairs
    const int* excl_start;   // n_colblocks + 1: entries of row block b are [excl_start[b], excl_start[b + 1])
    const int* excl_col;     // per entry: its column block (>= the row block)
    const unsigned long long* excl_mask;  // per entry, 64 rows: bit c = pair (row, column c) excluded
    int* flags_all;          // OR of every system's flags
    double* bbox_part;       // S x kBboxMaxSlices x 10: per slice of a system's atoms min, max, sum (xyz), bad
    int* bbox_done;          // S: slices of the system whose bounding box is in bbox_part
    int bbox_slices;         // the largest slice count of a system (bbox_slices())
    const int* atom_sys;     // N
    const int* cell_sys;     // cell_total
    // inputs
    const double* pos;       // N x 3
    const double* charge;    // N
    const double* radius;    // N
    const double* reach;     // N: tight grid reach R' + w + margin
    const float4* aparam;    // N: GBSW (R', (R'-w)^2, (R'+w)^2, R); GBMV1 (gamma, 0, 0, R);
                             //    GBMV2 (C, C - 100R^2, 400C^2, R)
    const double* quad_off;  // 3 x rows (rows = sum of the systems' Q)
    const double* quad_pw;   // 2 x rows: w_q / r_q^(order_k - 2)
    const float* quad_r2;    // rows: |offset_q|^2
    int quad_rows;
    const int4* pair_tiles;  // (row begin, column begin, column end, piece), global pair slots
    const int* pair_atom;    // N: the atom of each pair slot, a permutation inside each system (identity,
                             // or for cutoff systems a serpentine order so 64-slot blocks are compact)
    const int* colblk;       // n_colblocks: first atom of every 64-atom block
    // grid
    int* cell_fill;          // cell_total
    int* cell_start;         // cell_total + 1
    int* cell_count_eff;     // cell_total
    int* cell_atoms_tmp;     // entry_cap
    int* cell_atoms;         // entry_cap
    float4* ent_pos;         // entry_cap: position relative to the cell corner, .w = atom index bits
    float4* ent_par;         // entry_cap: aparam of the atom
    // points
    int* cell_pt_count;      // cell_total: points per cell (gathered from cell_ctr by the scan)
    int* cell_ctr;           // ctr_len: the binning passes' point counters, cell c's at ctr_slot(c)
                             // (ctr_perm = 1), or cell_pt_count itself (ctr_perm = 0)
    int ctr_perm;
    int ctr_len;             // cell_total rounded up to whole tiles of 4,096 (ctr_perm = 1)
    int* brick_count;        // brick_total x kBrickStride: point slots reserved in each brick (one counter per line)
    uint2* brick_rec;        // per system: the prec slot range, brick b owning [b sk^3 pt_cap, (b + 1) sk^3 pt_cap):
                             // (packed point, cell of the point inside the brick)
    int brick_total;         // sum of the systems' brick ranges
    int brick_max;           // the largest brick range of a system
    int any_bricks;          // some system bins by bricks (SysDesc::sk > 1)
    int any_cells;           // some system bins by cells (SysDesc::sk = 1)
    int* cell_act_count;     // cell_total
    int* frec_base;          // cell_total + 1: exclusive scan of the point counts
    int* worklist;           // VOL_KINDS x cell_total
    int* worklist_sys;       // VOL_KINDS x cell_total: system of each worklist cell
    float4* prec;            // per system: cell_cap x pt_cap slots; records (FP32 point - corner, packed point),
                             // or, binning by bricks, packed points (atom << kQBits | q) in the range's first quarter
    float4* frec0;           // NQ: active points, grouped by cell (corner-relative FP32 point, packed point)
    float4* frec1;           // NQ: GBSW (V, 0, 0, 0) / GBMV1 (dV/dS, V, 0, 0) / GBMV2 (c1, c2, c3, vx)
    float4* frec2;           // NQ: GBMV2 (vy, vz, V, 0)
    float* vval;             // NQ by pid: V, sign bit set on the points with a chain-rule factor
    float4* vrec;            // NQ by pid: owner force of those points only (others stale), .w unused
    float* vsig;             // NQ (adjoint, GBMV) by pid: S of the sigmoid, NaN where V has no derivative
    // per atom
    double* integrals;       // 2 x N
    double* born;            // N
    double* dbdi;            // 2 x N
    double* dedi;            // 2 x N
    double* pair_part;       // max_splits x 6 x N (fx, fy, fz, dE/dB, E, dE/dq): row sums per tile
    float* colpart;          // 5 planes x colpart_total (fx, fy, fz, dE/dB, dE/dq) per (row block, atom)
    double* atom_energy;     // N
    double* dedq;            // N (adjoint)
    long long* dedr_fix;     // N (adjoint): sum_p s fac_p dV_p/dR_j (GBSW: dR'_j), 2^32
    double* dedr;            // N (adjoint): dE/dR
    long long* vacc_fix;     // S x kVaccCopies x kVaccStride (adjoint), 2^32: copy c of system s's VA_COUNT
                             // sums at (s kVaccCopies + c) kVaccStride
    double* theta_atom;      // TH_COUNT x N (adjoint): per-atom parts of dE/dtheta
    double* theta;           // S x TH_COUNT (adjoint): dE/dtheta
    long long* force_fix;    // 3 x N, scale 2^32
    double* energy;          // S
    int* work;               // counters [8]
    int* scan_tmp;           // scan block sums
};

// Launchers (gbi_kernels.hip).  All enqueue on `stream`.
// marks: nullptr or kPasses + 1 events recorded before each pass and after the last.
// kinds: which volume kinds occur in the batch.
constexpr int kPasses = 8;
hipError_t launch_evaluate(const KParams& p, const bool* kinds, hipStream_t stream, const hipEvent_t* marks);
hipError_t launch_stats(const KParams& p, double* out6, hipStream_t stream);

// OpenMM GPU-platform interop (gbi_set_positions_posq, gbi_add_forces_fixed).
hipError_t launch_gather_posq(double* pos, const void* posq, int posq_double, const int* atom_slot, int n,
                              hipStream_t stream);
hipError_t launch_add_forces_fixed(long long* dst, int padded_n, const long long* force_fix, const int* atom_slot,
                                   int n, hipStream_t stream);

// gbi_execute_noalloc: bytes of the static device scratch array (per GPU), and its device
// address on the current device (gbi_noalloc.hip).
constexpr size_t kNoAllocScratchBytes = size_t(256) << 20;
hipError_t noalloc_scratch(void** ptr);

}  // namespace gbi
