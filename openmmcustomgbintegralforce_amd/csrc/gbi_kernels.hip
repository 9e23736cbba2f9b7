// gfx950 kernels of the integral Generalized-Born engine.
//
// One evaluation (launch_evaluate) of a batch of systems runs, on one stream, one launch
// per pass for all systems together:
//   clear -> bounding boxes / grid geometry (workgroup per system) -> cell lists (count,
//   scan, fill, sort+cap) -> point binning -> volume pass (per cell, lanes = points, one
//   launch per volume kind present) -> integrals + Born map (per atom) -> STILL pair pass
//   (tiles of 64 rows x a column range inside one system) -> per-atom chain (dE/dB ->
//   dE/dI, self energy) -> energy per system -> chain-rule force pass (per cell, lanes =
//   list atoms x points) -> owner-force reduction [-> adjoint reductions].
//
// Determinism: every sum is taken in a fixed order, except the per-(cell, atom) force
// partials, which are quantised to 2^-24 and summed exactly in FP64 before one int64
// fixed-point atomic per (cell, atom); integer addition is associative, so forces are
// bitwise reproducible.
//
// Reference counterparts: platforms/cuda/src/kernels/lookupTable.cu (grid),
// computeIntegral.cu + compute{GBSW,GBMV1,GBMV2}Volume.cu (volume pass),
// customGBValuePerParticle.cu (Born map), customGBEnergyN2.cu (pair pass),
// customGBEnergyPerParticle.cu (dE/dB -> dE/dI), reduceIntegralForce.cu +
// compute{GBSW,GBMV1,GBMV2}Force.cu (chain-rule pass).  The adjoint passes have no
// reference counterpart.
#include "gbi_internal.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <climits>
#include <algorithm>
#include <cmath>

namespace gbi {

namespace {

constexpr int kWave = 64;
constexpr int kScanThreads = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr double kFix32 = 4294967296.0;  // 2^32
// s_waitcnt immediate for vmcnt(0) alone (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// 6:4 and lgkmcnt 11:8 at their maxima, i.e. not waited for)
constexpr int kWaitVmem = 0x0F70;

__device__ inline double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ inline double wave_min(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ inline double wave_max(double v) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ inline int wave_scan_incl(int v, int lane) {
    for (int off = 1; off < kWave; off <<= 1) {
        int t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

// Any system of the batch hit a fatal condition: the host grows its buffers and redoes
// the whole evaluation, so every later pass stops early.
__device__ inline bool batch_fatal(const KParams& p) { return (*p.flags_all & kFatalFlags) != 0; }

__device__ inline void raise_flag(const KParams& p, int s, int flag) {
    atomicOr(&p.geom[s].flags, flag);
    atomicOr(p.flags_all, flag);
}

__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Exact sum of quantised partials: round x to the 2^-24 grid; the double holds an integer.
__device__ inline double quant(float x) { return (double)rintf(x * kQuantum); }

// Largest |value| the int64 2^-32 fixed point holds, in 2^-24 units and as a plain number
// (2^62 / 2^8 and 2^62 / 2^32: room for the sums of a few such values).
constexpr double kFixLimit24 = 1.8e16;
constexpr double kFixLimit = 1.0e9;

// 2^-24 units (exact integers in a double) -> int64 in 2^-32 units.
__device__ inline void add_fix(long long* dst, double units24) {
    atomicAdd((unsigned long long*)dst, (unsigned long long)(long long)(units24 * 256.0));
}

// ---------------------------------------------------------------------------
// Grid geometry and membership (FP64, no contraction: bitwise the oracle's arithmetic)
// ---------------------------------------------------------------------------

// floor(fl(a / g)) for 0 <= a / g < 2^20 without a division where it cannot matter:
// t = a * fl(1/g) differs from fl(a / g) by at most 3.3 ulp(a / g) < 1e-9, so floor(t) is
// the same integer unless t lies within 1e-9 of an integer; there the division decides.
__device__ inline int floor_div(double a, double g, double inv_g) {
#pragma clang fp contract(off)
    const double t = a * inv_g;
    const double f = floor(t);
    const double fr = t - f;
    if (fr > 1e-9 && fr < 1.0 - 1e-9) return (int)f;
    return (int)floor(a / g);
}

// Periodic grid: the point wrapped into the box (written back) and its cell, as
// CustomGBIntegral.cpp:294-301: x0 = x - floor((x - min)/L) L, index floor((x0 - min)/step).
__device__ inline int point_cell_periodic(const SysGeom& G, double* pt, int* idx) {
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
        const double pw = pt[d] - floor((pt[d] - G.origin[d]) / G.L[d]) * G.L[d];
        int k = (int)floor((pw - G.origin[d]) / G.step[d]);
        k = k < 0 ? 0 : (k > G.n[d] - 1 ? G.n[d] - 1 : k);
        idx[d] = k;
        pt[d] = pw;
    }
    return (idx[0] * G.n[1] + idx[1]) * G.n[2] + idx[2];
}

// x wrapped into the periodic box [origin, origin + L) (the cell-list build and its entries).
__device__ inline double wrap_coord(const SysGeom& G, int d, double x) {
#pragma clang fp contract(off)
    return x - G.L[d] * floor((x - G.origin[d]) / G.L[d]);
}

// Local cell of a point (and its index per axis); -1 outside.  Faithful: CustomGBIntegral.cpp:310-324.
__device__ inline int point_cell(const SysDesc& sd, const SysGeom& G, double px, double py, double pz, int* idx) {
#pragma clang fp contract(off)
    if (sd.grid_kind == GRID_BRUTE) {
        idx[0] = idx[1] = idx[2] = 0;
        return 0;
    }
    const double pt[3] = {px, py, pz};
    if (sd.grid_kind == GRID_FAITHFUL) {
        for (int d = 0; d < 3; ++d) {
            if (pt[d] < G.origin[d] || pt[d] > G.max[d]) return -1;
            idx[d] = floor_div(pt[d] - G.origin[d], sd.g, sd.inv_g);
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            double t = (pt[d] - G.origin[d]) * sd.inv_g;
            if (!(t >= 0.0) || t >= (double)G.n[d]) return -1;
            idx[d] = (int)floor(t);
            if (idx[d] >= G.n[d]) return -1;
        }
    }
    return (idx[0] * G.n[1] + idx[1]) * G.n[2] + idx[2];
}

// Minimum corner of the cell with indices (ix, iy, iz).
__device__ inline void cell_corner_idx(const SysDesc& sd, const SysGeom& G, int ix, int iy, int iz, double* corner) {
#pragma clang fp contract(off)
    if (sd.grid_kind == GRID_BRUTE) {
        corner[0] = G.origin[0];
        corner[1] = G.origin[1];
        corner[2] = G.origin[2];
        return;
    }
    if (sd.periodic) {
        corner[0] = G.origin[0] + ix * G.step[0];
        corner[1] = G.origin[1] + iy * G.step[1];
        corner[2] = G.origin[2] + iz * G.step[2];
        return;
    }
    corner[0] = G.origin[0] + ix * sd.g;
    corner[1] = G.origin[1] + iy * sd.g;
    corner[2] = G.origin[2] + iz * sd.g;
}

__device__ inline void cell_corner(const SysDesc& sd, const SysGeom& G, int c, double* corner) {
    const int nyz = G.n[1] * G.n[2];
    const int ix = c / nyz;
    const int rem = c - ix * nyz;
    const int iy = rem / G.n[2];
    const int iz = rem - iy * G.n[2];
    cell_corner_idx(sd, G, ix, iy, iz, corner);
}

// ---------------------------------------------------------------------------
// Clear and geometry
// ---------------------------------------------------------------------------

__global__ void k_clear(KParams p) {
    const int stride = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.cell_total; i += stride) {
        p.cell_fill[i] = 0;
        p.cell_ctr[i] = 0;
    }
    for (int i = p.cell_total + blockIdx.x * blockDim.x + threadIdx.x; i < p.ctr_len; i += stride) p.cell_ctr[i] = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.brick_total; i += stride)
        p.brick_count[(long long)i * kBrickStride] = 0;
    if (p.adjoint)
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < p.S * kVaccCopies * kVaccStride; i += stride)
            p.vacc_fix[i] = 0;
    if (blockIdx.x == 0) {
        if (threadIdx.x < 8) p.work[threadIdx.x] = 0;
        if (threadIdx.x == 0) *p.flags_all = 0;
        for (int k = threadIdx.x; k < p.S; k += blockDim.x) p.bbox_done[k] = 0;
    }
}

__device__ void bbox_to_geom(const KParams& p, int s, const double* mn, const double* mx, const double* sum, int isbad);

// Bounding box and centre of geometry -> grid geometry.  Grid (slices, S): workgroup b of
// system s reduces slice b of its atoms (bbox_slices(n) slices of ceil(n / slices) atoms) to a
// partial box; the workgroup that finishes a system's last slice combines the partials in
// slice order (a fixed order: the centre's sum does not depend on which finishes last) and
// derives the grid.  One workgroup per system took 0.012 ms for 5,000 atoms and 0.025 ms for
// 25,000 (a serial chain of position loads).
__global__ void __launch_bounds__(256) k_bbox(KParams p) {
    __shared__ double smin[3][4], smax[3][4], ssum[3][4];
    __shared__ int sbad[4];
    __shared__ int s_last;
    const int s = blockIdx.y, b = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int ab = p.sys[s].atom_begin, ae = p.sys[s].atom_end;
    const int nsl = bbox_slices(ae - ab);
    if (b >= nsl) return;  // the system needs fewer slices than the largest one
    const int per = (ae - ab + nsl - 1) / nsl;
    const int sb = ab + b * per, se = min(ae, sb + per);
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX}, sm[3] = {0.0, 0.0, 0.0};
    int bad = 0;
#pragma unroll 4
    for (int a = sb + tid; a < se; a += blockDim.x) {
        for (int d = 0; d < 3; ++d) {
            const double v = p.pos[3 * a + d];
            if (!isfinite(v)) bad = 1;
            mn[d] = fmin(mn[d], v);
            mx[d] = fmax(mx[d], v);
            sm[d] += v;
        }
    }
    for (int d = 0; d < 3; ++d) {
        mn[d] = wave_min(mn[d]);
        mx[d] = wave_max(mx[d]);
        sm[d] = wave_sum(sm[d]);
    }
    const int anybad = __any(bad);
    if (lane == 0) {
        for (int d = 0; d < 3; ++d) {
            smin[d][wid] = mn[d];
            smax[d][wid] = mx[d];
            ssum[d][wid] = sm[d];
        }
        sbad[wid] = anybad;
    }
    __syncthreads();
    if (tid == 0) {
        double* part = p.bbox_part + ((long long)s * kBboxMaxSlices + b) * kBboxPart;
        for (int d = 0; d < 3; ++d) {
            double a0 = smin[d][0], a1 = smax[d][0], a2 = ssum[d][0];
            for (int w = 1; w < 4; ++w) {
                a0 = fmin(a0, smin[d][w]);
                a1 = fmax(a1, smax[d][w]);
                a2 += ssum[d][w];
            }
            part[d] = a0;
            part[3 + d] = a1;
            part[6 + d] = a2;
        }
        part[9] = (double)(sbad[0] | sbad[1] | sbad[2] | sbad[3]);
        __threadfence();  // the partial is visible before the count that announces it
        s_last = atomicAdd(&p.bbox_done[s], 1) == nsl - 1;
    }
    __syncthreads();
    if (!s_last || tid != 0) return;
    __threadfence();  // every other slice's partial is visible after its count
    double gmn[3], gmx[3], gsm[3];
    int gbad = 0;
    for (int k = 0; k < nsl; ++k) {
        const double* part = p.bbox_part + ((long long)s * kBboxMaxSlices + k) * kBboxPart;
        for (int d = 0; d < 3; ++d) {
            gmn[d] = k ? fmin(gmn[d], part[d]) : part[d];
            gmx[d] = k ? fmax(gmx[d], part[3 + d]) : part[3 + d];
            gsm[d] = k ? gsm[d] + part[6 + d] : part[6 + d];
        }
        gbad |= part[9] != 0.0;
    }
    bbox_to_geom(p, s, gmn, gmx, gsm, gbad);
}

// Grid geometry from the system's bounding box and coordinate sum.
__device__ void bbox_to_geom(const KParams& p, int s, const double* mn, const double* mx, const double* sum, int isbad) {
#pragma clang fp contract(off)
    const SysDesc& sd = p.sys[s];
    double cog[3];
    for (int d = 0; d < 3; ++d) cog[d] = sum[d] / (double)(sd.atom_end - sd.atom_begin);
    SysGeom G;
    G.flags = 0;
    G.nb[0] = G.nb[1] = G.nb[2] = 0;
    G.halfdiag = sd.halfdiag;
    for (int d = 0; d < 3; ++d) {
        G.step[d] = sd.g;
        G.L[d] = 0.0;
    }
    long long ncell = 1;
    if (isbad) {
        G.flags |= FLAG_NONFINITE_POS;
        ncell = 0;
        for (int d = 0; d < 3; ++d) {
            G.origin[d] = 0.0;
            G.max[d] = 0.0;
            G.n[d] = 0;
        }
    } else if (sd.periodic) {
        // CustomGBIntegral.cpp:53-107: origin at the centre of geometry minus half the box,
        // ceil(L/g) + 1 cells of edge L/n per axis
        double maxstep = 0.0;
        for (int d = 0; d < 3; ++d) {
            const double L = p.box[3 * s + d];
            const double cells = ceil(L / sd.g) + 1.0;
            if (!(cells <= 1.0e6)) G.flags |= FLAG_GRID_TOO_LARGE;
            const int n = cells <= 1.0e6 ? (int)cells : 1;
            G.n[d] = n;
            G.L[d] = L;
            G.step[d] = L / n;
            G.origin[d] = cog[d] - L / 2.0;
            G.max[d] = G.origin[d] + L;
            maxstep = fmax(maxstep, G.step[d]);
            ncell *= n;
        }
        G.halfdiag = sqrt(3.0) / 2.0 * maxstep;
    } else if (sd.grid_kind == GRID_BRUTE) {
        for (int d = 0; d < 3; ++d) {
            G.origin[d] = 0.5 * (mn[d] + mx[d]);
            G.max[d] = DBL_MAX;
            G.n[d] = 1;
        }
    } else if (sd.grid_kind == GRID_FAITHFUL) {
        // CustomGBIntegral.cpp:173-187
        for (int d = 0; d < 3; ++d) {
            const double lo = mn[d] - sd.gpad;
            const double hi = mx[d] + sd.gpad;
            const double len = hi - lo;
            if (len > 1000.0) G.flags |= FLAG_GRID_TOO_LARGE;
            const int n = len > 1000.0 ? 1 : (int)ceil(len / sd.g) + 1;
            G.n[d] = n;
            G.origin[d] = lo;
            G.max[d] = lo + (n - 1) * sd.g;
            ncell *= n;
        }
    } else {
        for (int d = 0; d < 3; ++d) {
            const double lo = mn[d] - sd.gpad;
            const double hi = mx[d] + sd.gpad;
            const double len = hi - lo;
            if (len > 1000.0) G.flags |= FLAG_GRID_TOO_LARGE;
            const int n = len > 1000.0 ? 1 : (int)floor(len * sd.inv_g) + 1;
            G.n[d] = n;
            G.origin[d] = lo;
            G.max[d] = lo + n * sd.g;
            ncell *= n;
        }
    }
    // the cells of whole bricks (sk cells per axis, the last one cut at the grid's edge) must fit
    // the system's range: brick b's point slots are those of cells [b sk^3, (b + 1) sk^3)
    long long padded = 1;
    for (int d = 0; d < 3; ++d) {
        G.nb[d] = (G.n[d] + sd.sk - 1) / sd.sk;
        padded *= (long long)G.nb[d] * sd.sk;
    }
    if (ncell > sd.cell_cap || padded > sd.cell_cap) {
        G.flags |= FLAG_CELL_OVERFLOW;
        ncell = 0;
    }
    if (G.flags & kFatalFlags) ncell = 0;
    G.ncell = (int)ncell;
    p.geom[s] = G;
    if (G.flags) atomicOr(p.flags_all, G.flags);
    // the brute cell holds every atom of the system (its count is known here)
    if (sd.grid_kind == GRID_BRUTE && ncell == 1) p.cell_fill[sd.cell_base] = sd.atom_end - sd.atom_begin;
}

// ---------------------------------------------------------------------------
// Cell lists
// ---------------------------------------------------------------------------

// Candidate cells of atom j (before the membership test), clamped to the grid.
struct CandBox {
    int b[3], e[3];
    int count;
};

__device__ inline CandBox candidate_box(const SysDesc& sd, const SysGeom& G, const double* x, double pad) {
#pragma clang fp contract(off)
    CandBox B;
    if (sd.periodic) {
        // unwrapped cell indices around the wrapped atom; each is reduced modulo n later
        for (int d = 0; d < 3; ++d) {
            B.b[d] = (int)floor((x[d] - pad - G.origin[d]) / G.step[d]);
            B.e[d] = sd.grid_kind == GRID_FAITHFUL ? (int)ceil((x[d] + pad - G.origin[d]) / G.step[d])
                                                   : (int)floor((x[d] + pad - G.origin[d]) / G.step[d]);
        }
        const int nx = B.e[0] - B.b[0] + 1, ny = B.e[1] - B.b[1] + 1, nz = B.e[2] - B.b[2] + 1;
        B.count = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;
        return B;
    }
    for (int d = 0; d < 3; ++d) {
        if (sd.grid_kind == GRID_FAITHFUL) {
            B.b[d] = (int)floor((x[d] - pad - G.origin[d]) / sd.g);
            B.e[d] = (int)ceil((x[d] + pad - G.origin[d]) / sd.g);
        } else {
            B.b[d] = (int)floor((x[d] - pad - G.origin[d]) * sd.inv_g);
            B.e[d] = (int)floor((x[d] + pad - G.origin[d]) * sd.inv_g);
        }
        if (B.b[d] < 0) B.b[d] = 0;
        if (B.e[d] > G.n[d] - 1) B.e[d] = G.n[d] - 1;
    }
    const int nx = B.e[0] - B.b[0] + 1, ny = B.e[1] - B.b[1] + 1, nz = B.e[2] - B.b[2] + 1;
    B.count = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;
    return B;
}

// Membership of atom j in cell (i, jj, k).
// Faithful: corner within pad_j = R_j + sqrt3/2 g + buffer (CustomGBIntegral.cpp:210-223).
// Tight: distance from x_j to the cell box below reach_j = R'_j + w + margin.
__device__ inline bool cell_member(const SysDesc& sd, const SysGeom& G, const double* x, double pad, int i, int jj,
                                   int k) {
#pragma clang fp contract(off)
    if (sd.periodic) {
        // the cell image (i, jj, k) of the periodic grid; x wrapped into the box
        const double lx = G.origin[0] + i * G.step[0], ly = G.origin[1] + jj * G.step[1],
                     lz = G.origin[2] + k * G.step[2];
        if (sd.grid_kind == GRID_FAITHFUL) {
            const double dx = lx - x[0], dy = ly - x[1], dz = lz - x[2];
            return dx * dx + dy * dy + dz * dz < pad * pad;
        }
        const double dx = fmax(fmax(lx - x[0], x[0] - (lx + G.step[0])), 0.0);
        const double dy = fmax(fmax(ly - x[1], x[1] - (ly + G.step[1])), 0.0);
        const double dz = fmax(fmax(lz - x[2], x[2] - (lz + G.step[2])), 0.0);
        return dx * dx + dy * dy + dz * dz < pad * pad;
    }
    const double g = sd.g;
    if (sd.grid_kind == GRID_FAITHFUL) {
        const double dx = (G.origin[0] + i * g) - x[0];
        const double dy = (G.origin[1] + jj * g) - x[1];
        const double dz = (G.origin[2] + k * g) - x[2];
        return dx * dx + dy * dy + dz * dz < pad * pad;
    }
    const double lx = G.origin[0] + i * g, ly = G.origin[1] + jj * g, lz = G.origin[2] + k * g;
    const double dx = fmax(fmax(lx - x[0], x[0] - (lx + g)), 0.0);
    const double dy = fmax(fmax(ly - x[1], x[1] - (ly + g)), 0.0);
    const double dz = fmax(fmax(lz - x[2], x[2] - (lz + g)), 0.0);
    return dx * dx + dy * dy + dz * dz < pad * pad;
}

__device__ inline double atom_pad(const KParams& p, const SysDesc& sd, const SysGeom& G, int j) {
#pragma clang fp contract(off)
    if (sd.grid_kind != GRID_FAITHFUL) return p.reach[j];
    return (p.radius[j] + (sd.periodic ? G.halfdiag : sd.halfdiag)) + sd.buffer;
}

// Cell-list key: atom index << 6 | image code; the code holds, per axis in two bits, the
// box vector (-1, 0, +1) between the wrapped atom and the cell image it joined (0 when
// not periodic).  Keys sort by atom index first.
constexpr int kImageBits = 6;

// i = img n + iw with 0 <= iw < n (img is -1, 0 or +1 for boxes the host accepts; the
// floor division keeps every index inside the grid whatever i is).
__device__ inline int wrap_index(int i, int n, int& img) {
    img = i >= 0 ? i / n : -((n - 1 - i) / n);
    return i - img * n;
}

// The atom and cell of candidate t of box B (unwrapped indices reduced for periodic grids).
__device__ inline bool candidate_cell(const SysDesc& sd, const SysGeom& G, const CandBox& B, int t, const double* x,
                                      double pad, int& cell, int& code) {
    const int ny = B.e[1] - B.b[1] + 1, nz = B.e[2] - B.b[2] + 1;
    const int i = B.b[0] + t / (ny * nz);
    const int r = t - (i - B.b[0]) * (ny * nz);
    const int jj = B.b[1] + r / nz;
    const int k = B.b[2] + (r - (jj - B.b[1]) * nz);
    if (!cell_member(sd, G, x, pad, i, jj, k)) return false;
    if (sd.periodic) {
        int ox, oy, oz;
        const int iw = wrap_index(i, G.n[0], ox), jw = wrap_index(jj, G.n[1], oy), kw = wrap_index(k, G.n[2], oz);
        cell = (iw * G.n[1] + jw) * G.n[2] + kw;
        code = (ox & 3) | ((oy & 3) << 2) | ((oz & 3) << 4);
    } else {
        cell = (i * G.n[1] + jj) * G.n[2] + k;
        code = 0;
    }
    return true;
}

// The atom position the lists use: wrapped into the box for periodic systems.
__device__ inline void list_position(const KParams& p, const SysDesc& sd, const SysGeom& G, int j, double* x) {
    for (int d = 0; d < 3; ++d) x[d] = sd.periodic ? wrap_coord(G, d, p.pos[3 * j + d]) : p.pos[3 * j + d];
}

// Appends key to cell's raw list (fixed capacity per cell; the atomic's count stays exact on
// overflow, so the host can size the lists from it).
__device__ inline void append_member(const KParams& p, const SysDesc& sd, int s, int cell, int key) {
    const int slot = atomicAdd(&p.cell_fill[sd.cell_base + cell], 1);
    if (slot < sd.tmp_cap)
        p.cell_atoms_tmp[sd.tmp_base + (long long)cell * sd.tmp_cap + slot] = key;
    else
        raise_flag(p, s, FLAG_LIST_OVERFLOW);
}

constexpr int kMaxWalk = 1024;  // samples per axis of the periodic walk (about 12 at the reference's sizes)

// The reference's periodic membership of CharmmGBMVForce grids (CustomGBIntegral.cpp:114-157):
// sample points walked from x - pad - g' to x + pad + g' (g' = max L/n) in steps of L/n per
// axis, each coordinate accumulated in FP64 as the reference's loop does (sample k of an axis
// = k successive additions); a sample within pad of the atom (x unwrapped) puts the atom in
// the cell of the sample wrapped into the box, with the image code of that wrap.  One lane
// per sample.
__device__ void fill_periodic_samples(const KParams& p, const SysDesc& sd, int s, const SysGeom& G, int j) {
#pragma clang fp contract(off)
    const double x[3] = {p.pos[3 * j], p.pos[3 * j + 1], p.pos[3 * j + 2]};
    const double pad = atom_pad(p, sd, G, j);
    const double gmax = fmax(fmax(G.step[0], G.step[1]), G.step[2]);
    double bc[3], ec[3];
    int cnt[3], kx[3];
    for (int d = 0; d < 3; ++d) {
        bc[d] = x[d] - pad - gmax;
        ec[d] = x[d] + pad + gmax;
        int k = 0;
        for (double v = bc[d]; v < ec[d] && k < kMaxWalk; v += G.step[d]) ++k;
        cnt[d] = k;
        kx[d] = (int)floor((x[d] - G.origin[d]) / G.L[d]);  // the wrap of the atom (wrap_coord)
    }
    if (cnt[0] >= kMaxWalk || cnt[1] >= kMaxWalk || cnt[2] >= kMaxWalk) {  // a radius of many boxes
        if (threadIdx.x == 0) raise_flag(p, s, FLAG_GRID_TOO_LARGE);
        return;
    }
    const long long total = (long long)cnt[0] * cnt[1] * cnt[2];
    for (long long t = threadIdx.x; t < total; t += blockDim.x) {
        const int i3[3] = {(int)(t / (cnt[1] * cnt[2])), (int)((t / cnt[2]) % cnt[1]), (int)(t % cnt[2])};
        double smp[3];
        for (int d = 0; d < 3; ++d) {
            double v = bc[d];
            for (int k = 0; k < i3[d]; ++k) v += G.step[d];
            smp[d] = v;
        }
        const double dx = smp[0] - x[0], dy = smp[1] - x[1], dz = smp[2] - x[2];
        if (!(dx * dx + dy * dy + dz * dz < pad * pad)) continue;
        int idx[3], code = 0;
        for (int d = 0; d < 3; ++d) {
            const double ks = floor((smp[d] - G.origin[d]) / G.L[d]);
            const double w = smp[d] - G.L[d] * ks;
            int k = (int)floor((w - G.origin[d]) / G.step[d]);
            idx[d] = k < 0 ? 0 : (k > G.n[d] - 1 ? G.n[d] - 1 : k);
            code |= (((int)ks - kx[d]) & 3) << (2 * d);  // the atom's image next to the cell
        }
        append_member(p, sd, s, (idx[0] * G.n[1] + idx[1]) * G.n[2] + idx[2], (j << kImageBits) | code);
    }
}

// One workgroup per atom, lanes over its candidate cells: each member key goes to its cell's
// raw list.  The CSR layout of the sorted lists is the scan of the counts.  Atoms in the
// binning order (systems interleaved, each strided), so the workgroups in flight append to
// different systems' and distant cells' counters.
__global__ void __launch_bounds__(256) k_grid_fill(KParams p) {
    if (batch_fatal(p)) return;
    const int j = p.bin_order[blockIdx.x];
    const int s = p.atom_sys[j];
    const SysDesc sd = p.sys[s];  // a copy: read once, not again after each append
    if (sd.grid_kind == GRID_BRUTE) {
        // index order, no atomics (k_bbox set the count)
        if (threadIdx.x == 0) {
            if (j - sd.atom_begin < sd.tmp_cap)
                p.cell_atoms_tmp[sd.tmp_base + (j - sd.atom_begin)] = j << kImageBits;
            else
                raise_flag(p, s, FLAG_LIST_OVERFLOW);
        }
        return;
    }
    const SysGeom G = p.geom[s];
    if (sd.periodic && sd.grid_kind == GRID_FAITHFUL) {
        fill_periodic_samples(p, sd, s, G, j);
        return;
    }
    double x[3];
    list_position(p, sd, G, j, x);
    const double pad = atom_pad(p, sd, G, j);
    const CandBox B = candidate_box(sd, G, x, pad);
    for (int t = threadIdx.x; t < B.count; t += blockDim.x) {
        int cell, code;
        if (candidate_cell(sd, G, B, t, x, pad, cell, code)) append_member(p, sd, s, cell, (j << kImageBits) | code);
    }
}

// Entry of key (atom, image code) relative to the cell corner: for periodic systems the
// wrapped atom moved by the code's box vectors, i.e. its image nearest the cell.
__device__ inline void write_entry(const KParams& p, const SysDesc& sd, const SysGeom& G, int e, int key,
                                   const double* corner) {
#pragma clang fp contract(off)
    const int j = key >> kImageBits;
    double x[3];
    for (int d = 0; d < 3; ++d) {
        x[d] = p.pos[3 * j + d];
        if (sd.periodic) {
            const int bits = (key >> (2 * d)) & 3;
            const int img = bits == 3 ? -1 : bits;
            x[d] = wrap_coord(G, d, x[d]) - img * G.L[d];  // the cell image sits img boxes away
        }
    }
    p.ent_pos[e] = make_float4((float)(x[0] - corner[0]), (float)(x[1] - corner[1]), (float)(x[2] - corner[2]),
                               __int_as_float(j));
    p.ent_par[e] = p.aparam[j];
}

// One wave per cell: order the cell's atoms by index (rank sort), keep the first `cap`
// (the reference appends in ascending atom order and stops at the cap,
// CustomGBIntegral.cpp:219-222) and write their corner-relative entries.  (Ordering GBSW
// lists by distance to the cell, to reach V = 0 sooner, was measured slower.)  Only the
// periodic sample walk can list an atom twice in a cell (two samples of its walk in the
// cell), as the reference would; equal keys then take consecutive ranks.
constexpr int kSortLds = 256;  // cells with more raw members sort in global memory

__global__ void __launch_bounds__(64) k_grid_sort(KParams p) {
    __shared__ __align__(16) int s_in[kSortLds];
    __shared__ int s_out[kSortLds];
    if (batch_fatal(p)) return;
    const int lane = threadIdx.x;
    for (int c = blockIdx.x; c < p.cell_total; c += gridDim.x) {
        const int n_raw = p.cell_fill[c];
        if (n_raw == 0) {
            if (lane == 0) p.cell_count_eff[c] = 0;
            continue;
        }
        const int s = p.cell_sys[c];
        const SysDesc sd = p.sys[s];  // a copy: read once, not again after each entry's stores
        const int n = min(n_raw, sd.tmp_cap);  // (more is a fatal flag: this evaluation is redone)
        const SysGeom G = p.geom[s];
        const int base = p.cell_start[c];
        const int eff = n < sd.cap ? n : sd.cap;
        const int* raw = p.cell_atoms_tmp + sd.tmp_base + (long long)(c - sd.cell_base) * sd.tmp_cap;
        double corner[3];
        cell_corner(sd, G, c - sd.cell_base, corner);
        if (sd.grid_kind == GRID_BRUTE) {
            for (int k = lane; k < eff; k += kWave) write_entry(p, sd, G, base + k, raw[k], corner);
        } else if (n <= kSortLds) {
            // keys are distinct; padded to a multiple of 4 with keys above them all, so each
            // rank reads the keys 4 at a time
            const int n4 = (n + 3) & ~3;
            for (int k = lane; k < n4; k += kWave) s_in[k] = k < n ? raw[k] : INT_MAX;
            __syncthreads();
            const bool ties = sd.periodic && sd.grid_kind == GRID_FAITHFUL;
            for (int k = lane; k < n; k += kWave) {
                const int a = s_in[k];
                int r = 0;
                for (int m = 0; m < n4; m += 4) {
                    const int4 q = *reinterpret_cast<const int4*>(&s_in[m]);
                    r += (q.x < a) + (q.y < a) + (q.z < a) + (q.w < a);
                }
                if (ties)
                    for (int m = 0; m < k; ++m) r += s_in[m] == a;
                s_out[r] = a;
            }
            __syncthreads();
            for (int k = lane; k < eff; k += kWave) write_entry(p, sd, G, base + k, s_out[k], corner);
        } else {
            for (int k = lane; k < n; k += kWave) {
                const int a = raw[k];
                int r = 0;
                for (int m = 0; m < n; ++m) r += raw[m] < a || (raw[m] == a && m < k);
                p.cell_atoms[base + r] = a;
            }
            __threadfence_block();
            __syncthreads();
            for (int k = lane; k < eff; k += kWave) write_entry(p, sd, G, base + k, p.cell_atoms[base + k], corner);
        }
        if (lane == 0) p.cell_count_eff[c] = eff;
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Exclusive scan of int arrays (two passes, deterministic)
// ---------------------------------------------------------------------------

// Slot of cell c's binning counter: within each tile of 4,096 counters, cell r of the tile at
// (r mod 128) x 32 + r / 128, so the 32 counters of a 128-byte line belong to cells 128 apart
// and the cells of one atom's points, mostly neighbours, reserve on different lines.  Dense:
// the clear and the scan's gather read whole lines.
__host__ __device__ inline long long ctr_slot(long long c) {
    return (c & ~4095ll) | ((c & 127) << 5) | ((c >> 7) & 31);
}
__device__ inline long long cell_ctr_slot(const KParams& p, long long c) { return p.ctr_perm ? ctr_slot(c) : c; }

// perm: the inputs are the binning counters in[ctr_slot(idx)], copied to compact[idx] on the way.
__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int* in, int* out, int n, int* block_sums,
                                                             int perm, int* compact) {
    __shared__ int s_wave[kScanThreads / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const long long base = (long long)blockIdx.x * kScanTile + (long long)tid * kScanItems;
    int v[kScanItems];
    int tsum = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const long long idx = base + k;
        v[k] = idx < n ? in[perm ? ctr_slot(idx) : idx] : 0;
        if (perm && idx < n) compact[idx] = v[k];
        tsum += v[k];
    }
    const int incl = wave_scan_incl(tsum, lane);
    if (lane == kWave - 1) s_wave[wid] = incl;
    __syncthreads();
    if (wid == 0) {
        const int nw = kScanThreads / kWave;
        int w = lane < nw ? s_wave[lane] : 0;
        const int wi = wave_scan_incl(w, lane);
        if (lane < nw) s_wave[lane] = wi - w;
        if (lane == nw - 1) block_sums[blockIdx.x] = wi;
    }
    __syncthreads();
    int run = s_wave[wid] + incl - tsum;
    for (int k = 0; k < kScanItems; ++k) {
        const long long idx = base + k;
        if (idx < n) out[idx] = run;
        run += v[k];
    }
}

// Adds to tile t the sum of the tile sums before it (each workgroup sums them itself: at
// most kScanTile ints, exact in any order); the last tile also writes the total.
// Sum of the tile sums before tile t, by a whole workgroup of kScanThreads threads (every
// thread gets it; s_wave is reused on return).
__device__ inline int tiles_before(const int* block_sums, int t, int* s_wave) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    int part = 0;
    for (int k = tid; k < t; k += kScanThreads) part += block_sums[k];
    const int wsum = wave_scan_incl(part, lane);
    if (lane == kWave - 1) s_wave[wid] = wsum;
    __syncthreads();
    int add = 0;
    for (int w = 0; w < kScanThreads / kWave; ++w) add += s_wave[w];
    __syncthreads();
    return add;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_add(int* out, int n, const int* block_sums, int limit,
                                                           int* flags, int flag) {
    __shared__ int s_wave[kScanThreads / kWave];
    const int tid = threadIdx.x;
    const int t = blockIdx.x;
    const int add = tiles_before(block_sums, t, s_wave);
    if (t == (int)gridDim.x - 1 && tid == 0) {
        const int total = add + block_sums[t];
        out[n] = total;
        if (flags && total > limit) atomicOr(flags, flag);
    }
    const long long base = (long long)t * kScanTile;
    for (int k = tid; k < kScanTile; k += kScanThreads) {
        const long long idx = base + k;
        if (idx < n) out[idx] += add;
    }
}

// out[0..n] = exclusive scan of in[0..n-1] (out[n] = total); flags |= flag when total > limit.
// perm: the inputs are in[ctr_slot(k)], also copied to compact[k].
hipError_t scan_exclusive(const int* in, int* out, int n, int* tmp, int limit, int* flags, int flag,
                          hipStream_t stream, int perm = 0, int* compact = nullptr) {
    const int nblocks = (n + kScanTile - 1) / kScanTile;
    if (nblocks > kScanTile || nblocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_scan_tiles, dim3(nblocks), dim3(kScanThreads), 0, stream, in, out, n, tmp, perm, compact);
    hipLaunchKernelGGL(k_scan_add, dim3(nblocks), dim3(kScanThreads), 0, stream, out, n, (const int*)tmp, limit, flags,
                       flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Point binning
// ---------------------------------------------------------------------------

// r_q = x_i + off_q in FP64 (the reference's point, GBSWIntegral.cpp:132-133).
__device__ inline void quad_point(const KParams& p, double xi, double yi, double zi, int row, double* pt) {
#pragma clang fp contract(off)
    pt[0] = xi + p.quad_off[3 * row];
    pt[1] = yi + p.quad_off[3 * row + 1];
    pt[2] = zi + p.quad_off[3 * row + 2];
}

// Binning.  k_point_bin: one workgroup per atom (systems in turn, each visited with a stride
// coprime to its size so concurrent workgroups hit distant counters) finds the cell of every
// quadrature point (FP64, the oracle's arithmetic), counts its points per key in an LDS hash
// table (local ranks) and reserves each key's slots with one global atomic; these returning
// atomics bound the pass (3.4 M on per-cell counters took 0.135 ms of 0.19 for 5,000 atoms).
// Two ways, per system:
//   bricks (sk > 1): the key is the brick of sk x sk x sk cells holding the point's cell, its
//     counter on a 128-byte line of its own; the workgroup files (packed point, cell in brick)
//     in the brick's slots and k_point_split moves the packed points into the cells.  An atom's 1,200 GBMV-II
//     points fall into about 200 bricks of 0.45 nm against 670 cells of 0.15 nm.
//   cells (sk = 1, the exact GBSW grid of 0.20 nm, where bricks of 0.40 nm save too few
//     atomics to pay for the second pass): the key is the cell and the workgroup writes the
//     cell's slots directly.
// What a cell slot holds depends on the way (pt_slot / point_record):
//   bricks: the packed point (atom << kQBits | q), 4 bytes; the volume pass recomputes the
//     point from the atom's FP64 position and the quadrature row.  Against a 16-byte record
//     this halves k_point_split (0.062 -> 0.030 ms at 5,000 GBMV-II atoms) for 0.005 ms more
//     in the volume pass.
//   cells: the record (the point relative to its cell's corner in FP32, the packed point),
//     16 bytes.  There the binning pass is bound by its atomics, not by its writes, and the
//     GBSW volume pass, whose points mostly leave after a few atoms on burial, took 20 % longer
//     recomputing the points.
// Points outside the grid, in an empty cell, or (GBSW) inside the owner's own burial radius get
// their volume directly.
constexpr int kBinHash = 2048;  // > Q: the table never fills
// Threads per workgroup: bricks 512 (4 points per thread, 56 VGPRs, 8 waves per SIMD: 0.088 ->
// 0.075 ms at 5,000 GBMV-II atoms against 256 threads); cells 256 (8 points per thread: the
// 512-thread form took 0.62 ms against 0.42 at 25,000 GBSW atoms, whose pass is bound by its
// returning atomics).
template <bool BRICKS>
constexpr int bin_threads() { return BRICKS ? 512 : 256; }

// Corner-relative FP32 record of point row of atom i in the cell with minimum corner `corner`:
// the FP64 point as binned (r_q = x_i + off_q, wrapped into a periodic box) minus the corner,
// and the packed point.
__device__ inline float4 point_record(const KParams& p, const SysDesc& sd, const SysGeom& G, int i, int row,
                                      uint32_t packed, const double* corner) {
#pragma clang fp contract(off)
    double pt[3];
    quad_point(p, p.pos[3 * i], p.pos[3 * i + 1], p.pos[3 * i + 2], row, pt);
    if (sd.periodic) {
        int idx[3];
        (void)point_cell_periodic(G, pt, idx);
    }
    return make_float4((float)(pt[0] - corner[0]), (float)(pt[1] - corner[1]), (float)(pt[2] - corner[2]),
                       __uint_as_float(packed));
}

// Slot s of local cell c of a system binning by bricks: the packed points live in the first
// quarter of the system's own float4 slot range.
__device__ inline uint32_t* pt_slot(const KParams& p, const SysDesc& sd, int c, int s) {
    return reinterpret_cast<uint32_t*>(p.prec) + 4 * sd.prec_base + (long long)c * sd.pt_cap + s;
}

// BRICKS: the launch for the systems binning by bricks (the other systems' workgroups return).
// Bricks: 8 waves per SIMD (the system's fields held in scalar registers would otherwise
// cost one); cells: register bound at 5.
template <bool BRICKS>
__global__ void __launch_bounds__(bin_threads<BRICKS>(), BRICKS ? 8 : 1) k_point_bin(KParams p) {
    constexpr int kBinThreads = bin_threads<BRICKS>();
    constexpr int kBinPer = kBinHash / kBinThreads;  // points per thread
    // t_cnt: the points per key, then the key's first reserved slot
    __shared__ int t_key[kBinHash], t_cnt[kBinHash];
    if (batch_fatal(p)) return;
    const int tid = threadIdx.x;
    const int i = p.bin_order[blockIdx.x];
    const int s = p.atom_sys[i];
    // copies: the fields are read once, not again per point (the LDS atomics between points
    // keep the compiler from reusing values loaded through a reference)
    const SysDesc sd = p.sys[s];
    const SysGeom G = p.geom[s];
    const int Q = sd.Q, qb = sd.quad_base, cb = sd.cell_base, sk = sd.sk, pt_cap = sd.pt_cap;
    constexpr bool bricks = BRICKS;
    if ((sk > 1) != BRICKS) return;
    const long long key_slots = (long long)sk * sk * sk * pt_cap;
    const long long prec_base = sd.prec_base;
    const int pbase = sd.pt_base + (i - sd.atom_begin) * Q;
    // only lists that surely hold the owner (exact or brute) may skip it; the reference
    // GBMV-form grid is under-padded and can miss it (GBSWIntegral.cpp:59-60)
    const bool shortcut = sd.vol_kind == VOL_GBSW && sd.grid_kind != GRID_FAITHFUL;
    const float v_empty = sd.v_empty;
    const float sig_empty = sd.vol_kind == VOL_GBMV1 ? 0.0f : __int_as_float(0x7fc00000);
    const double xi = p.pos[3 * i], yi = p.pos[3 * i + 1], zi = p.pos[3 * i + 2];
    const float own_lo2 = p.aparam[i].y;
    for (int k = tid; k < kBinHash; k += kBinThreads) {
        t_key[k] = -1;
        t_cnt[k] = 0;
    }
    __syncthreads();
    int h[kBinPer], rank[kBinPer], cell[kBinPer];       // bricks: the point's cell in its brick
    float rx[kBinPer], ry[kBinPer], rz[kBinPer];        // cells: the record's coordinates
#pragma unroll
    for (int u = 0; u < kBinPer; ++u) {
        h[u] = -1;
        rank[u] = 0;
        cell[u] = 0;
        rx[u] = ry[u] = rz[u] = 0.f;
        const int q = tid + kBinThreads * u;
        if (q >= Q) continue;
        const int row = qb + q;
        int c = -1, idx[3] = {0, 0, 0};
        double pt[3] = {0.0, 0.0, 0.0};
        const bool owner_buried = shortcut && p.quad_r2[row] <= own_lo2;
        if (!owner_buried) {
            quad_point(p, xi, yi, zi, row, pt);
            c = sd.periodic ? point_cell_periodic(G, pt, idx) : point_cell(sd, G, pt[0], pt[1], pt[2], idx);
            if (c >= 0 && p.cell_count_eff[cb + c] == 0) c = -1;
        }
        if (c < 0) {
            p.vval[pbase + q] = owner_buried ? 0.0f : v_empty;
            if (p.adjoint) p.vsig[pbase + q] = owner_buried ? __int_as_float(0x7fc00000) : sig_empty;
            continue;
        }
        int key = c;
        if constexpr (bricks) {
            const int bx = idx[0] / sk, by = idx[1] / sk, bz = idx[2] / sk;
            key = (bx * G.nb[1] + by) * G.nb[2] + bz;
            cell[u] = ((idx[0] - bx * sk) * sk + (idx[1] - by * sk)) * sk + (idx[2] - bz * sk);
        } else {
            double corner[3];  // point_record's arithmetic, pt already wrapped
            cell_corner_idx(sd, G, idx[0], idx[1], idx[2], corner);
            rx[u] = (float)(pt[0] - corner[0]);
            ry[u] = (float)(pt[1] - corner[1]);
            rz[u] = (float)(pt[2] - corner[2]);
        }
        // open addressing; at most Q < kBinHash distinct keys, so a free entry always exists
        int e = (int)(((unsigned)key * 2654435761u) >> 21);
        for (int probe = 0; probe < kBinHash; ++probe) {
            const int old = atomicCAS(&t_key[e], -1, key);
            if (old == -1 || old == key) break;
            e = (e + 1) & (kBinHash - 1);
        }
        h[u] = e;
        rank[u] = atomicAdd(&t_cnt[e], 1);
    }
    __syncthreads();
    for (int k = tid; k < kBinHash; k += kBinThreads) {
        const int n = t_cnt[k];
        if (n > 0) {
            if constexpr (bricks)
                t_cnt[k] = atomicAdd(&p.brick_count[(long long)(sd.brick_base + t_key[k]) * kBrickStride], n);
            else
                t_cnt[k] = atomicAdd(&p.cell_ctr[cell_ctr_slot(p, cb + t_key[k])], n);
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kBinPer; ++u) {
        if (h[u] < 0) continue;
        const int q = tid + kBinThreads * u;
        const long long slot = (long long)t_cnt[h[u]] + rank[u];
        const int key = t_key[h[u]];
        const uint32_t packed = ((uint32_t)i << kQBits) | (uint32_t)q;
        if (slot >= key_slots) {
            raise_flag(p, s, FLAG_POINT_OVERFLOW);  // bricks: then some cell of the brick has more than pt_cap
        } else if constexpr (bricks) {
            p.brick_rec[prec_base + key * key_slots + slot] = make_uint2(packed, (uint32_t)cell[u]);
        } else {
            p.prec[prec_base + (long long)key * pt_cap + slot] = make_float4(rx[u], ry[u], rz[u], __uint_as_float(packed));
        }
    }
}

// k_point_split moves each brick's points into their cells.  A workgroup takes one chunk of
// kSplitChunk points of a brick at a time (blockIdx.x: brick, blockIdx.z: chunk, blockIdx.y:
// system; the fullest bricks hold 4x the mean, so whole bricks per workgroup would leave a long
// tail), sorts the chunk's packed points by cell in LDS, reserves each cell's slots with one
// global atomic (at most sk^3 per chunk) and writes the packed points in that order.
constexpr int kSplitChunk = 2048;
constexpr int kSplitPer = kSplitChunk / 256;
constexpr int kSplitParts = 4;  // workgroups per brick (each takes every 4th chunk)

__global__ void __launch_bounds__(256) k_point_split(KParams p) {
    constexpr int kCells = kMaxBrickEdge * kMaxBrickEdge * kMaxBrickEdge;
    __shared__ uint2 s_rec[kSplitChunk];
    __shared__ int s_cnt[kCells], s_off[kCells], s_base[kCells];
    if (batch_fatal(p)) return;
    const int s = blockIdx.y, tid = threadIdx.x;
    const SysDesc sd = p.sys[s];  // copies, read once (as in k_point_bin)
    const SysGeom G = p.geom[s];
    const int sk = sd.sk, sk3 = sk * sk * sk, cb = sd.cell_base, pt_cap = sd.pt_cap;
    const int nbrick = sk > 1 ? G.nb[0] * G.nb[1] * G.nb[2] : 0;  // cells binned directly: nothing to move
    const long long brick_slots = (long long)sk3 * pt_cap;
    for (int b = blockIdx.x; b < nbrick; b += gridDim.x) {
        const int n = (int)min((long long)p.brick_count[(long long)(sd.brick_base + b) * kBrickStride], brick_slots);
        const int bz = b % G.nb[2], by = (b / G.nb[2]) % G.nb[1], bx = b / (G.nb[1] * G.nb[2]);
        const uint2* rec = p.brick_rec + sd.prec_base + b * brick_slots;
        // n is uniform over the workgroup, so every thread takes the same chunks
        for (int k0 = blockIdx.z * kSplitChunk; k0 < n; k0 += gridDim.z * kSplitChunk) {
            const int m = min(kSplitChunk, n - k0);
            __syncthreads();
            if (tid < sk3) s_cnt[tid] = 0;
            __syncthreads();
            uint2 r[kSplitPer];  // (packed point, its cell in the brick)
            int rank[kSplitPer];
#pragma unroll
            for (int j = 0; j < kSplitPer; ++j) {
                const int k = tid + 256 * j;
                r[j] = k < m ? rec[k0 + k] : make_uint2(0u, 0xffffffffu);
            }
#pragma unroll
            for (int j = 0; j < kSplitPer; ++j) rank[j] = r[j].y == 0xffffffffu ? -1 : atomicAdd(&s_cnt[r[j].y], 1);
            __syncthreads();
            if (tid < kWave) {  // the first wave: exclusive scan of the cell counts, slot reservations
                const int cnt = tid < sk3 ? s_cnt[tid] : 0;
                int incl = cnt;
#pragma unroll
                for (int d = 1; d < kWave; d <<= 1) {
                    const int t = __shfl_up(incl, d, kWave);
                    if (tid >= d) incl += t;
                }
                if (tid < sk3) s_off[tid] = incl - cnt;
                if (cnt > 0) {
                    const int ix = bx * sk + tid / (sk * sk), iy = by * sk + (tid / sk) % sk, iz = bz * sk + tid % sk;
                    const int c = cb + (ix * G.n[1] + iy) * G.n[2] + iz;
                    s_base[tid] = atomicAdd(&p.cell_ctr[cell_ctr_slot(p, c)], cnt);
                }
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kSplitPer; ++j)
                if (rank[j] >= 0) s_rec[s_off[r[j].y] + rank[j]] = r[j];
            __syncthreads();
            for (int k = tid; k < m; k += 256) {
                const uint2 e = s_rec[k];
                const int u = (int)e.y;
                const int slot = s_base[u] + (k - s_off[u]);
                if (slot >= pt_cap) {
                    raise_flag(p, s, FLAG_POINT_OVERFLOW);
                    continue;
                }
                const int ix = bx * sk + u / (sk * sk), iy = by * sk + (u / sk) % sk, iz = bz * sk + u % sk;
                const int c = (ix * G.n[1] + iy) * G.n[2] + iz;
                *pt_slot(p, sd, c, slot) = e.x;
            }
        }
    }
}

// Cells holding points, one list per volume kind (order irrelevant: every later sum is
// order independent).  One atomic per workgroup and kind: all of them go to one counter
// per kind, which serialises them (one per wave took 0.009 ms for 770 waves at 5,000 atoms,
// one per 1,024-thread workgroup 0.002 ms; 0.014 -> 0.005 ms at 25,000).
// It also finishes the exclusive scan of the point counts (frec_base), whose tile pass ran
// before it: each chunk of kWorkThreads cells lies in one scan tile and adds the sum of the
// tile sums before it (the work of k_scan_add, without a launch of its own).
constexpr int kWorkThreads = 1024;
static_assert(kWorkThreads == kScanThreads && kScanTile % kWorkThreads == 0, "a chunk lies in one scan tile");

__global__ void __launch_bounds__(kWorkThreads) k_worklist(KParams p) {
    constexpr int kWaves = kWorkThreads / kWave;
    __shared__ int s_cnt[VOL_KINDS][kWaves];
    __shared__ int s_base[VOL_KINDS];
    __shared__ int s_scan[kScanThreads / kWave];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int stride = gridDim.x * blockDim.x;
    const int ntiles = (p.cell_total + kScanTile - 1) / kScanTile;
    for (int base = blockIdx.x * blockDim.x; base < p.cell_total; base += stride) {
        const int t = base / kScanTile;
        const int add = tiles_before(p.scan_tmp, t, s_scan);
        const int c = base + (int)threadIdx.x;
        if (c < p.cell_total) p.frec_base[c] += add;
        if (c == p.cell_total - 1) {  // the last cell: the total and its limit
            const int total = add + p.scan_tmp[ntiles - 1];
            p.frec_base[p.cell_total] = total;
            if (total > p.NQ) atomicOr(p.flags_all, FLAG_POINT_OVERFLOW);
        }
    }
    if (batch_fatal(p)) return;
    for (int base = blockIdx.x * blockDim.x; base < p.cell_total; base += stride) {
        const int c = base + (int)threadIdx.x;
        int vk = -1, s = 0;
        if (c < p.cell_total && p.cell_pt_count[c] > 0) {
            s = p.cell_sys[c];
            vk = p.sys[s].vol_kind;
        }
        unsigned long long m[VOL_KINDS];
#pragma unroll
        for (int k = 0; k < VOL_KINDS; ++k) {
            m[k] = __ballot(vk == k);
            if (lane == 0) s_cnt[k][wid] = __popcll(m[k]);
        }
        __syncthreads();
        if (threadIdx.x < VOL_KINDS) {
            int total = 0;
            for (int w = 0; w < kWaves; ++w) total += s_cnt[threadIdx.x][w];
            s_base[threadIdx.x] = total > 0 ? atomicAdd(&p.work[threadIdx.x], total) : 0;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < VOL_KINDS; ++k) {
            if (vk != k) continue;
            int off = s_base[k];
            for (int w = 0; w < wid; ++w) off += s_cnt[k][w];
            const int slot = k * p.cell_total + off + __popcll(m[k] & ((1ull << lane) - 1ull));
            p.worklist[slot] = c;
            p.worklist_sys[slot] = s;
        }
        __syncthreads();  // s_cnt and s_base are reused by the next iteration
    }
}

// ---------------------------------------------------------------------------
// Volume pass: one wave per cell, lanes = quadrature points, the cell's atoms staged in
// LDS (corner-relative FP32).  Writes V per point (vval, by point; stored negated, -0 for
// V = 0, on the points that carry a chain-rule factor) and their compacted records.
// GBSW: V = prod_j H_j, early exit on burial (GBSWIntegral.cpp:187-224);
// GBMV-I: S = sum exp(gamma_j d^4) (GBMVIntegralTypeI.cpp:155-173);
// GBMV-II: sum1, sum2, v, S = S0 sum1 sum2/|v|^2 (GBMVIntegralTypeII.cpp:163-191).
// ---------------------------------------------------------------------------

// The sigmoid V = 1/(1 + exp(t)), t = beta (S - lambda), of GBMV-I/II is taken as its limit
// (V = 1 for t < 0, V = 0 for t > 0, no chain-rule term) where |t| > kSigmoidSat: exp(t) then
// underflows or overflows FP32, |1 - V| or V is below 2e-38 and dV/dt = -V (1 - V) below
// 2e-38 as well, so the FP64 sigmoid agrees to that level whatever beta and lambda are.
constexpr float kSigmoidSat = 87.0f;

// V = 1/(1 + exp(t)) and V (1 - V) = exp(t)/(1 + exp(t))^2, from e = exp(-|t|) <= 1 so that
// nothing overflows for |t| up to kSigmoidSat (exp(t) (1 + exp(t))^-2 written directly
// overflows FP32 from t = 44 on).
__device__ inline void sigmoid(float t, float& V, float& vv1) {
    const float e = __expf(-fabsf(t));
    const float ie = 1.f / (1.f + e);
    V = t <= 0.f ? ie : e * ie;
    vv1 = e * ie * ie;
}

// V >= 0 with the sign bit as the "has a chain-rule factor" mark (also set for V = 0)
__device__ inline float mark_active(float v, bool active) {
    return __int_as_float(__float_as_int(v) | (active ? (int)0x80000000 : 0));
}
__device__ inline float volume_of(float vv) { return __int_as_float(__float_as_int(vv) & 0x7fffffff); }

template <int VK>
__global__ void __launch_bounds__(64, 7) k_volume(KParams p) {
    __shared__ float s_x[kMaxAtomChunk], s_y[kMaxAtomChunk], s_z[kMaxAtomChunk];
    __shared__ float s_a[kMaxAtomChunk], s_b[kMaxAtomChunk], s_c[kMaxAtomChunk];
    if (batch_fatal(p)) return;
    const int lane = threadIdx.x;
    const int nwork = p.work[VK];
    const int* worklist = p.worklist + VK * p.cell_total;
    const int* worklist_sys = p.worklist_sys + VK * p.cell_total;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int c = worklist[w];
        const int sys = worklist_sys[w];
        // references (copies put part of the geometry in scratch memory: the GBMV-II pass
        // took 0.22 ms instead of 0.16 at 5,000 atoms)
        const SysDesc& sd = p.sys[sys];
        const SysGeom& G = p.geom[sys];
        const int pt_cap = sd.pt_cap;
        const int np = min(p.cell_pt_count[c], pt_cap);
        const int na = p.cell_count_eff[c];
        const long long ps = sd.prec_base + (long long)(c - sd.cell_base) * pt_cap;
        const int as = p.cell_start[c];
        const bool single = na <= kMaxAtomChunk;
        const int fb = p.frec_base[c];  // exclusive scan of the point counts
        const int pt_base = sd.pt_base, abeg = sd.atom_begin, Q = sd.Q, qb = sd.quad_base;
        const bool ids = sd.sk > 1;  // the slots hold packed points (bricks), else records
        double corner[3];
        cell_corner(sd, G, c - sd.cell_base, corner);
        const float k1 = sd.k1, k3 = sd.k3, fclamp = sd.fclamp, S0 = sd.S0, vguard = sd.vguard;
        const float beta = sd.beta, lambda = sd.lambda;
        const bool acc64 = VK != VOL_GBSW && sd.acc64;
        auto stage = [&](int ab, int cnt) {
            for (int k = lane; k < cnt; k += kWave) {
                const float4 e = p.ent_pos[as + ab + k];
                const float4 q = p.ent_par[as + ab + k];
                s_x[k] = e.x;
                s_y[k] = e.y;
                s_z[k] = e.z;
                s_a[k] = q.x;
                s_b[k] = q.y;
                s_c[k] = q.z;
            }
            if (VK == VOL_GBSW && lane < ((cnt + 3) & ~3) - cnt) {
                // pad to a multiple of 4 with atoms that neither bury nor touch any point
                const int k = cnt + lane;
                s_x[k] = s_y[k] = s_z[k] = s_a[k] = 0.f;
                s_b[k] = s_c[k] = -1.f;
            }
        };
        __syncthreads();
        if (single) stage(0, na);
        __syncthreads();
        int act = 0;
        for (int pb = 0; pb < np; pb += kWave) {
            const int k = pb + lane;
            const bool valid = k < np;
            float4 pr = make_float4(0.f, 0.f, 0.f, 0.f);
            uint32_t pk = 0u;
            if (valid) {
                if (ids) {
                    pk = *pt_slot(p, sd, c - sd.cell_base, k);
                    pr = point_record(p, sd, G, (int)(pk >> kQBits), qb + (int)(pk & ((1u << kQBits) - 1)), pk, corner);
                } else {
                    pr = p.prec[ps + k];
                    pk = __float_as_uint(pr.w);
                }
            }
            const int i = (int)(pk >> kQBits), q = (int)(pk & ((1u << kQBits) - 1));
            const float prx = pr.x, pry = pr.y, prz = pr.z;
            float V = 1.f;
            bool done = !valid;
            float s1 = 0.f, s2 = 0.f, vx = 0.f, vy = 0.f, vz = 0.f;
            double S1 = 0.0, S2 = 0.0, VX = 0.0, VY = 0.0, VZ = 0.0;  // acc64 systems
            for (int ab = 0; ab < na; ab += kMaxAtomChunk) {
                const int cnt = min(kMaxAtomChunk, na - ab);
                if (!single) {
                    __syncthreads();
                    stage(ab, cnt);
                    __syncthreads();
                }
                if (VK == VOL_GBSW) {
                    // branch-free: H = 0 inside R' - w, the cubic in the shell, 1 beyond R' + w;
                    // the wave leaves once every point is buried (V = 0), checked every 4 atoms
                    if (done) V = 0.f;
                    const int cnt4 = (cnt + 3) & ~3;
                    for (int m = 0; m < cnt4; m += 4) {
                        if (__all(V == 0.f)) break;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const float dx = s_x[m + u] - prx, dy = s_y[m + u] - pry, dz = s_z[m + u] - prz;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            const float x = fast_sqrt(d2) - s_a[m + u];
                            float h = fmaxf(fmaf(x, fmaf(-k3 * x, x, k1), 0.5f), 0.f);
                            h = d2 < s_c[m + u] ? h : 1.f;
                            h = d2 <= s_b[m + u] ? 0.f : h;
                            V *= h;
                        }
                    }
                } else if (VK == VOL_GBMV2) {
                    if (valid && !acc64) {
                        #pragma unroll 4
                        for (int m = 0; m < cnt; ++m) {
                            const float dx = s_x[m] - prx, dy = s_y[m] - pry, dz = s_z[m] - prz;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            const float t = fmaf(100.f, d2, s_b[m]);
                            const float ct = s_a[m] * fast_rcp(t);
                            const float F = fminf(ct * ct, fclamp);
                            s1 += F;
                            s2 = fmaf(d2, F * F, s2);
                            vx = fmaf(F, dx, vx);
                            vy = fmaf(F, dy, vy);
                            vz = fmaf(F, dz, vz);
                        }
                    } else if (valid) {
                        // the same terms, summed in FP64
                        #pragma unroll 2
                        for (int m = 0; m < cnt; ++m) {
                            const float dx = s_x[m] - prx, dy = s_y[m] - pry, dz = s_z[m] - prz;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            const float t = fmaf(100.f, d2, s_b[m]);
                            const float ct = s_a[m] * fast_rcp(t);
                            const double F = (double)fminf(ct * ct, fclamp);
                            S1 += F;
                            S2 = fma((double)d2, F * F, S2);
                            VX = fma(F, (double)dx, VX);
                            VY = fma(F, (double)dy, VY);
                            VZ = fma(F, (double)dz, VZ);
                        }
                    }
                } else {
                    if (valid && !acc64) {
                        #pragma unroll 4
                        for (int m = 0; m < cnt; ++m) {
                            const float dx = s_x[m] - prx, dy = s_y[m] - pry, dz = s_z[m] - prz;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            s1 += __expf(s_a[m] * (d2 * d2));
                        }
                    } else if (valid) {
                        #pragma unroll 2
                        for (int m = 0; m < cnt; ++m) {
                            const float dx = s_x[m] - prx, dy = s_y[m] - pry, dz = s_z[m] - prz;
                            const float d2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            S1 += (double)__expf(s_a[m] * (d2 * d2));
                        }
                    }
                }
            }
            bool active = false;
            float4 r1 = make_float4(0.f, 0.f, 0.f, 0.f), r2 = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid) {
                const int pid = pt_base + (i - abeg) * Q + q;
                if (VK == VOL_GBSW) {
                    active = V > 0.f && V < 1.f;
                    p.vval[pid] = mark_active(V, active);
                    r1.x = V;
                } else if (VK == VOL_GBMV2) {
                    float vsq, S, t;
                    bool guard;
                    if (acc64) {
                        // S and t = beta (S - lambda) in FP64, then the FP32 record
                        const double vsq64 = fma(VX, VX, fma(VY, VY, VZ * VZ));
                        guard = !(vsq64 >= (double)vguard);
                        const double S64 = guard ? 0.0 : sd.S064 * S1 * S2 / vsq64;
                        S = (float)S64;
                        t = (float)(sd.beta64 * (S64 - sd.lambda64));
                        s1 = (float)S1;
                        s2 = (float)S2;
                        vx = (float)VX;
                        vy = (float)VY;
                        vz = (float)VZ;
                        vsq = (float)vsq64;
                    } else {
                        vsq = fmaf(vx, vx, fmaf(vy, vy, vz * vz));
                        // |v|^2 below the guard: the S -> +inf limit; otherwise the sigmoid's
                        // exponent t decides (its limit beyond +-kSigmoidSat)
                        guard = !(vsq >= vguard);
                        S = guard ? 0.f : S0 * s1 * s2 / vsq;
                        t = beta * (S - lambda);
                    }
                    const bool sat = guard || !(fabsf(t) <= kSigmoidSat);
                    float Vv = (guard ? beta < 0.f : t < 0.f) ? 1.f : 0.f;
                    if (!sat) {
                        float vv1;
                        sigmoid(t, Vv, vv1);
                        const float kk = -beta * vv1 * S0;
                        const float ivsq = 1.f / vsq;
                        r1 = make_float4(kk * s2 * ivsq, 2.f * kk * s1 * ivsq, -2.f * kk * s1 * s2 * ivsq * ivsq, vx);
                        r2 = make_float4(vy, vz, Vv, 0.f);
                        active = true;
                    }
                    p.vval[pid] = mark_active(Vv, active);
                    if (p.adjoint) p.vsig[pid] = sat ? __int_as_float(0x7fc00000) : S;
                } else {
                    if (acc64) s1 = (float)S1;
                    const float t = acc64 ? (float)(sd.beta64 * (S1 - sd.lambda64)) : beta * (s1 - lambda);
                    float Vv = t < 0.f ? 1.f : 0.f;
                    if (fabsf(t) <= kSigmoidSat) {
                        float vv1;
                        sigmoid(t, Vv, vv1);
                        r1.x = -beta * vv1;
                    }
                    r1.y = Vv;
                    active = r1.x != 0.f;
                    p.vval[pid] = mark_active(Vv, active);
                    if (p.adjoint) p.vsig[pid] = s1;
                }
            }
            const unsigned long long mask = __ballot(active);
            if (active) {
                const int o = fb + act + __popcll(mask & ((1ull << lane) - 1ull));
                p.frec0[o] = pr;
                p.frec1[o] = r1;
                if (VK == VOL_GBMV2) p.frec2[o] = r2;
            }
            act += __popcll(mask);
        }
        if (lane == 0) p.cell_act_count[c] = act;
    }
}

// ---------------------------------------------------------------------------
// Integrals and Born radii: one wave per atom, fixed-order FP64 sums over q.
// I_k(i) = sum_q w_q / r_q^(n_k - 2) * (1 - V) [GBSW] or * V [GBMV]
// (GBSWIntegral.cpp:142-146; GBMVIntegralTypeII.cpp:145-149).
// B = slope / (a0 (1/eta - I4) + a1 (1/(4 eta^4) - I7)^(1/4)) + shift
// (CharmmReferenceGBSW.cpp:591-593; dB/dI as :607-609).
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(64) k_integrals(KParams p) {
    if (batch_fatal(p)) return;
    const int i = blockIdx.x, lane = threadIdx.x;
    const int s = p.atom_sys[i];
    const SysDesc& sd = p.sys[s];
    const int Q = sd.Q, qb = sd.quad_base;
    const int pbase = sd.pt_base + (i - sd.atom_begin) * Q;
    double a0 = 0.0, a1 = 0.0;
    const bool gbsw = sd.vol_kind == VOL_GBSW;
    // Q a multiple of 4 (every system's points start on a multiple of 4): 4 points per lane
    // and load (0.045 -> 0.037 ms at 25k GBSW atoms), in a fixed order either way
    if ((Q & 3) == 0) {
        const float4* v4 = reinterpret_cast<const float4*>(p.vval + pbase);
        for (int q4 = lane; q4 < (Q >> 2); q4 += kWave) {
            const float4 w = v4[q4];
            const float vv[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int q = 4 * q4 + j;
                const double v = (double)volume_of(vv[j]);
                const double f = gbsw ? 1.0 - v : v;
                a0 += p.quad_pw[qb + q] * f;
                a1 += p.quad_pw[p.quad_rows + qb + q] * f;
            }
        }
    } else {
#pragma unroll 4
        for (int q = lane; q < Q; q += kWave) {
            const double v = (double)volume_of(p.vval[pbase + q]);
            const double f = gbsw ? 1.0 - v : v;
            a0 += p.quad_pw[qb + q] * f;
            a1 += p.quad_pw[p.quad_rows + qb + q] * f;
        }
    }
    a0 = wave_sum(a0);
    a1 = wave_sum(a1);
    if (lane != 0) return;
    p.integrals[i] = a0;
    p.integrals[p.N + i] = a1;
    if (sd.custom) return;  // the generated kernels map the integrals
    const double I4 = sd.i4 == 0 ? a0 : a1;
    const double I7 = sd.i7 == 0 ? a0 : a1;
    const double eta = sd.eta;
    const double eta4 = eta * eta * eta * eta;
    const double T = 1.0 / (4.0 * eta4) - I7;
    const double D = sd.alpha0 * (1.0 / eta - I4) + sd.alpha1 * pow(T, 0.25);
    const double B = sd.slope / D + sd.shift;
    p.born[i] = B;
    p.dbdi[sd.i4 * p.N + i] = sd.slope * sd.alpha0 / (D * D);
    p.dbdi[sd.i7 * p.N + i] = sd.slope * 0.25 * sd.alpha1 * pow(T, -0.75) / (D * D);
    if (!isfinite(B) || !(B > 0.0)) raise_flag(p, s, FLAG_NONFINITE_RESULT);
}

// ---------------------------------------------------------------------------
// STILL pair pass: E_ij = P q_i q_j / f_ij, f = sqrt(r^2 + B_i B_j exp(-r^2/(4 B_i B_j)))
// (CharmmReferenceGBSW.cpp:248-282).  A system's atoms form 64-atom blocks; a tile (one
// wave) is row block I with column blocks J0..J1-1, J >= I.  The diagonal block is
// evaluated from the row side (every ordered pair i != j); an off-diagonal block evaluates
// each unordered pair once: lane l holds row atom i and at step m meets column (l + m) mod
// 64, whose partial sums then move one lane down, so after 64 steps lane l holds column l's
// sums (fixed order, no atomics).  Row sums go to pair_part (one plane per tile of the row),
// column sums to colpart (per row block and atom); k_atom_chain adds them in a fixed order.
// Each block's terms are summed in FP32 (64 terms) and folded into FP64.
// ADJ: also sum_j P q_j / f_ij = dE/dq_i - P q_i/B_i, on both sides.
// ---------------------------------------------------------------------------

// One pair's terms (from i's side: d = x_j - x_i).  ex = exp(-r^2/(4 B_i B_j)) as
// exp2(-(r^2 kexp_i) / B_j) with kexp_i = log2(e) / (4 B_i); rf = 1/f.
struct PairTerms {
    float dx, dy, dz, r2, ex, rf;
};

__device__ inline PairTerms pair_terms(double xj, double yj, double zj, double xi, double yi, double zi, float Bi,
                                       float Bj, float kexp_i, float iBj) {
    PairTerms t;
    t.dx = (float)(xj - xi);
    t.dy = (float)(yj - yi);
    t.dz = (float)(zj - zi);
    t.r2 = fmaf(t.dx, t.dx, fmaf(t.dy, t.dy, t.dz * t.dz));
    t.ex = __builtin_amdgcn_exp2f(-(t.r2 * kexp_i) * iBj);
    t.rf = __builtin_amdgcn_rsqf(fmaf(Bi * Bj, t.ex, t.r2));
    return t;
}

// Cutoff systems.  FP64 squared distance of atoms i and j as the oracle takes it:
// d = x_j - x_i, reduced to the minimum image by d -= L floor(d/L + 1/2) (OpenMM's
// ReferenceForce::getDeltaRPeriodic), r^2 = (dx dx + dy dy) + dz dz.
__device__ inline double exact_r2(const KParams& p, int s, const SysDesc& sd, int i, int j) {
#pragma clang fp contract(off)
    double d[3];
    for (int k = 0; k < 3; ++k) {
        d[k] = p.pos[3 * j + k] - p.pos[3 * i + k];
        if (sd.periodic) d[k] -= p.box[3 * s + k] * floor(d[k] / p.box[3 * s + k] + 0.5);
    }
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
}

// Pair inside the cutoff: r <= c for the standalone GBSW force (CharmmReferenceGBSW.cpp:246-247,
// :290-291), r < c for CharmmGBMVForce (CharmmReferenceGBMV.cpp:300-301).  The FP32 distance
// decides outside a band of +-1e-4 c^2 around the cutoff, the FP64 one inside it, so the
// decision is the oracle's.
__device__ inline bool within_cutoff(const KParams& p, int s, const SysDesc& sd, float r2, int i, int j) {
    if (r2 < sd.cut2_lo) return true;
    if (r2 > sd.cut2_hi) return false;
    const double e = exact_r2(p, s, sd, i, j);
    return sd.cut_mode == 1 ? !(e > sd.cut2) : e < sd.cut2;
}

// Box vectors of a periodic system in FP32, for minimum images of short differences.
struct BoxF {
    float L[3], inv[3];
};

__device__ inline BoxF box_f(const KParams& p, int s, const SysDesc& sd) {
    BoxF b;
    for (int d = 0; d < 3; ++d) {
        b.L[d] = sd.periodic ? (float)p.box[3 * s + d] : 0.f;
        b.inv[d] = sd.periodic ? (float)(1.0 / p.box[3 * s + d]) : 0.f;
    }
    return b;
}

// d reduced to the minimum image (|d| below 1.5 box edges; L = 0: not periodic).
__device__ inline float min_image(float d, float L, float inv) { return fmaf(-L, rintf(d * inv), d); }

// A pair of a cutoff system: d at its minimum image when periodic (positions wrapped into the
// box, so one box vector at most).
__device__ inline PairTerms pair_terms_cut(double xj, double yj, double zj, double xi, double yi, double zi,
                                           bool per, const BoxF& bx, float Bi, float Bj, float kexp_i, float iBj) {
    PairTerms t;
    t.dx = (float)(xj - xi);
    t.dy = (float)(yj - yi);
    t.dz = (float)(zj - zi);
    if (per) {
        t.dx = min_image(t.dx, bx.L[0], bx.inv[0]);
        t.dy = min_image(t.dy, bx.L[1], bx.inv[1]);
        t.dz = min_image(t.dz, bx.L[2], bx.inv[2]);
    }
    t.r2 = fmaf(t.dx, t.dx, fmaf(t.dy, t.dy, t.dz * t.dz));
    t.ex = __builtin_amdgcn_exp2f(-(t.r2 * kexp_i) * iBj);
    t.rf = __builtin_amdgcn_rsqf(fmaf(Bi * Bj, t.ex, t.r2));
    return t;
}

__device__ inline double wrap_box(double x, double L) {
#pragma clang fp contract(off)
    return L > 0.0 ? x - L * floor(x / L) : x;
}

// Bounding box of every 64-atom block (positions wrapped into the box for periodic systems),
// for skipping pair blocks beyond the cutoff.
__global__ void __launch_bounds__(64) k_block_box(KParams p) {
    if (batch_fatal(p)) return;
    const int b = blockIdx.x, lane = threadIdx.x;
    const int j0 = p.colblk[b];
    const int s = p.atom_sys[j0];
    const SysDesc& sd = p.sys[s];
    const int j = p.pair_atom[min(j0 + lane, sd.atom_end - 1)];
    float lo[3], hi[3];
    for (int d = 0; d < 3; ++d) {
        const float v = (float)wrap_box(p.pos[3 * j + d], sd.periodic ? p.box[3 * s + d] : 0.0);
        lo[d] = (float)wave_min(v);
        hi[d] = (float)wave_max(v);
    }
    if (lane == 0) {
        p.blk_lo[b] = make_float4(lo[0], lo[1], lo[2], 0.f);
        p.blk_hi[b] = make_float4(hi[0], hi[1], hi[2], 0.f);
    }
}

// Blocks a and b surely farther apart than the cutoff (conservative: 0.1 % on c^2 and
// no test along an axis where the two boxes span half the periodic box).
__device__ inline bool blocks_apart(const KParams& p, const SysDesc& sd, const BoxF& bx, int a, int b) {
    const float4 la = p.blk_lo[a], ha = p.blk_hi[a], lb = p.blk_lo[b], hb = p.blk_hi[b];
    const float lo_a[3] = {la.x, la.y, la.z}, hi_a[3] = {ha.x, ha.y, ha.z};
    const float lo_b[3] = {lb.x, lb.y, lb.z}, hi_b[3] = {hb.x, hb.y, hb.z};
    float gap2 = 0.f;
    for (int d = 0; d < 3; ++d) {
        const float h = 0.5f * ((hi_a[d] - lo_a[d]) + (hi_b[d] - lo_b[d]));
        float dc = 0.5f * ((lo_b[d] + hi_b[d]) - (lo_a[d] + hi_a[d]));
        if (sd.periodic) {
            if (h >= 0.5f * bx.L[d]) continue;
            dc = min_image(dc, bx.L[d], bx.inv[d]);
        }
        const float g = fabsf(dc) - h;
        if (g > 0.f) gap2 = fmaf(g, g, gap2);
    }
    return gap2 > sd.cut2_hi * 1.001f;
}

// Column partial of row block I for local atom j (block J = j / 64 > I): triangular, the
// row blocks of one column block stored one after the other, 64 atoms each.
__device__ inline long long colpart_index(const SysDesc& sd, int I, int j) {
    const long long J = j / kWave;
    return sd.colpart_base + (long long)kWave * (J * (J - 1) / 2 + I) + (j - kWave * J);
}

// Lane l receives lane l + 1's value (mod 64): DPP wave_rol:1.
__device__ inline float rotate_down(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x134, 0xF, 0xF, false));
}

// CUT: systems with a cutoff drop the pairs outside it (rf = 0 zeroes every term), reduce
// periodic differences to the minimum image (positions wrapped into the box on load, so one
// box vector at most) and skip column blocks whose bounding boxes are beyond the cutoff
// (their column partials are written as zeros); excluded pairs (a ParticlePair term,
// CharmmReferenceGBMV.cpp:272-284) are dropped through the block's 64 x 64 bit mask.
template <bool ADJ, bool CUT>
__global__ void __launch_bounds__(64) k_pairs(KParams p) {
    __shared__ double s_px[64], s_py[64], s_pz[64];
    __shared__ float4 s_col[64];  // (q, B, 1/B, 1/(4B))
    __shared__ int s_atom[64];
    if (batch_fatal(p)) return;
    const int lane = threadIdx.x;
    const int4 tile = p.pair_tiles[blockIdx.x];
    const int s = p.atom_sys[tile.x];
    const SysDesc sd = p.sys[s];  // a copy: read once
    const int ab = sd.atom_begin, ae = sd.atom_end;
    const int i = tile.x + lane;  // pair slot; its atom is pair_atom[i]
    const bool iv = i < ae;
    const int ai = iv ? p.pair_atom[i] : 0;
    const bool cut = CUT && sd.cut_mode != 0;
    const bool per = CUT && sd.periodic;
    BoxF bx;
    double Lx = 0.0, Ly = 0.0, Lz = 0.0;
    if (CUT) {
        bx = box_f(p, s, sd);
        if (per) {
            Lx = p.box[3 * s];
            Ly = p.box[3 * s + 1];
            Lz = p.box[3 * s + 2];
        }
    }
    double xi = iv ? p.pos[3 * ai] : 0.0, yi = iv ? p.pos[3 * ai + 1] : 0.0, zi = iv ? p.pos[3 * ai + 2] : 0.0;
    if (per) {
        xi = wrap_box(xi, Lx);
        yi = wrap_box(yi, Ly);
        zi = wrap_box(zi, Lz);
    }
    const float P = (float)sd.prefactor;
    const float qi = iv ? (float)p.charge[ai] : 0.f;  // zero charge: an absent row adds nothing
    const float Bi = iv ? (float)p.born[ai] : 1.f;
    const float Pq = P * qi;
    const float q4Bi = 0.25f / Bi;
    const float kexp = 1.4426950408889634f * q4Bi;
    const int I_row = (tile.x - ab) / kWave;
    double fx = 0.0, fy = 0.0, fz = 0.0, dedb = 0.0, e = 0.0, dq = 0.0;
    for (int jb = tile.y; jb < tile.z; jb += kWave) {
        const int jj = jb + lane;
        const bool jv = jj < ae;
        if (cut && jb != tile.x && blocks_apart(p, sd, bx, sd.blk_base + I_row, sd.blk_base + (jb - ab) / kWave)) {
            if (jv) {
                const long long o = colpart_index(sd, I_row, jj - ab);
                const long long T = p.colpart_total;
                p.colpart[o] = 0.f;
                p.colpart[T + o] = 0.f;
                p.colpart[2 * T + o] = 0.f;
                p.colpart[3 * T + o] = 0.f;
                if (ADJ) p.colpart[4 * T + o] = 0.f;
            }
            continue;
        }
        unsigned long long emask = 0ull;  // bit c: pair (this row, column c of the block) excluded
        if (CUT && p.any_excl) {
            const int rb = sd.blk_base + I_row, cb = sd.blk_base + (jb - ab) / kWave;
            for (int e = p.excl_start[rb]; e < p.excl_start[rb + 1]; ++e)
                if (p.excl_col[e] == cb) {
                    emask = p.excl_mask[(size_t)e * kWave + lane];
                    break;
                }
        }
        __syncthreads();
        const int aj = jv ? p.pair_atom[jj] : 0;
        s_atom[lane] = aj;
        {
            double xj = jv ? p.pos[3 * aj] : 0.0, yj = jv ? p.pos[3 * aj + 1] : 0.0, zj = jv ? p.pos[3 * aj + 2] : 0.0;
            if (per) {
                xj = wrap_box(xj, Lx);
                yj = wrap_box(yj, Ly);
                zj = wrap_box(zj, Lz);
            }
            s_px[lane] = xj;
            s_py[lane] = yj;
            s_pz[lane] = zj;
        }
        {
            const float Bj = jv ? (float)p.born[aj] : 1.f;
            const float iB = 1.f / Bj;
            s_col[lane] = make_float4(jv ? (float)p.charge[aj] : 0.f, Bj, iB, 0.25f * iB);
        }
        __syncthreads();
        float efs = 0.f, dqs = 0.f, fxs = 0.f, fys = 0.f, fzs = 0.f, dbs = 0.f;
        if (jb == tile.x) {
            // diagonal block: ordered pairs from the row side, half the pair energy each
            const int cnt = min(kWave, ae - jb);
            for (int m = 0; m < cnt; ++m) {
                if (jb + m == i) continue;
                const float4 cj = s_col[m];
                PairTerms t;
                if (cut) {
                    t = pair_terms_cut(s_px[m], s_py[m], s_pz[m], xi, yi, zi, per, bx, Bi, cj.y, kexp, cj.z);
                    if (!iv || !within_cutoff(p, s, sd, t.r2, ai, s_atom[m])) t.rf = 0.f;
                } else {
                    t = pair_terms(s_px[m], s_py[m], s_pz[m], xi, yi, zi, Bi, cj.y, kexp, cj.z);
                }
                if (CUT && ((emask >> m) & 1ull)) t.rf = 0.f;
                const float eij = Pq * cj.x * t.rf;
                efs = fmaf(0.5f, eij, efs);
                if (ADJ) dqs = fmaf(P * cj.x, t.rf, dqs);
                const float t3 = eij * (t.rf * t.rf);
                const float dGdr = t3 * fmaf(0.25f, t.ex, -1.f);  // -Pqq (4 - ex) / (4 f^3)
                fxs = fmaf(t.dx, dGdr, fxs);
                fys = fmaf(t.dy, dGdr, fys);
                fzs = fmaf(t.dz, dGdr, fzs);
                dbs = fmaf(-0.5f * t3 * t.ex, fmaf(t.r2, q4Bi, cj.y), dbs);
            }
        } else {
            // off-diagonal block: each unordered pair once, column sums rotate through the lanes
            float cfx = 0.f, cfy = 0.f, cfz = 0.f, cdb = 0.f, cdq = 0.f;
            for (int m = 0; m < kWave; ++m) {
                const int c = (lane + m) & (kWave - 1);
                const float4 cj = s_col[c];
                PairTerms t;
                if (cut) {
                    t = pair_terms_cut(s_px[c], s_py[c], s_pz[c], xi, yi, zi, per, bx, Bi, cj.y, kexp, cj.z);
                    if (!iv || jb + c >= ae || !within_cutoff(p, s, sd, t.r2, ai, s_atom[c])) t.rf = 0.f;
                } else {
                    t = pair_terms(s_px[c], s_py[c], s_pz[c], xi, yi, zi, Bi, cj.y, kexp, cj.z);
                }
                if (CUT && ((emask >> c) & 1ull)) t.rf = 0.f;
                const float eij = Pq * cj.x * t.rf;  // P q_i q_j / f
                efs += eij;
                const float t3 = eij * (t.rf * t.rf);
                const float dGdr = t3 * fmaf(0.25f, t.ex, -1.f);  // -Pqq (4 - ex) / (4 f^3)
                const float hb = -0.5f * t3 * t.ex;               // dE/dB_i = hb (r^2/(4B_i) + B_j)
                fxs = fmaf(t.dx, dGdr, fxs);
                fys = fmaf(t.dy, dGdr, fys);
                fzs = fmaf(t.dz, dGdr, fzs);
                cfx = fmaf(-t.dx, dGdr, cfx);
                cfy = fmaf(-t.dy, dGdr, cfy);
                cfz = fmaf(-t.dz, dGdr, cfz);
                dbs = fmaf(hb, fmaf(t.r2, q4Bi, cj.y), dbs);
                cdb = fmaf(hb, fmaf(t.r2, cj.w, Bi), cdb);
                if (ADJ) {
                    dqs = fmaf(P * cj.x, t.rf, dqs);
                    cdq = fmaf(Pq, t.rf, cdq);
                }
                cfx = rotate_down(cfx);
                cfy = rotate_down(cfy);
                cfz = rotate_down(cfz);
                cdb = rotate_down(cdb);
                if (ADJ) cdq = rotate_down(cdq);
            }
            if (jv) {
                const long long o = colpart_index(sd, I_row, jj - ab);
                const long long T = p.colpart_total;
                p.colpart[o] = cfx;
                p.colpart[T + o] = cfy;
                p.colpart[2 * T + o] = cfz;
                p.colpart[3 * T + o] = cdb;
                if (ADJ) p.colpart[4 * T + o] = cdq;
            }
        }
        e += efs;
        fx += fxs;
        fy += fys;
        fz += fzs;
        dedb += dbs;
        if (ADJ) dq += dqs;
    }
    if (iv) {
        double* out = p.pair_part + (size_t)tile.w * 6 * p.N;
        out[i] = fx;
        out[p.N + i] = fy;
        out[2 * p.N + i] = fz;
        out[3 * p.N + i] = dedb;
        out[4 * p.N + i] = e;
        if (ADJ) out[5 * p.N + i] = dq;
    }
}

// Per atom, one workgroup per 64-atom block: the pair sums in a fixed order, then the self
// term and dE/dI = dE/dB dB/dI (customGBEnergyPerParticle.cu; CharmmReferenceGBSW.cpp:257-273).
// Thread (g, lane) sums, for atom lane of the block, the row pieces k = g, g + G, ... and the
// column partials of row blocks I = g, g + G, ... < bj in FP64 (G = 16 wave groups); the
// groups are added in order.  ADJ: dE/dq and the Born-map parts of dE/dtheta.
constexpr int kChainGroups = 16;

template <bool ADJ>
__global__ void __launch_bounds__(kChainGroups * kWave) k_atom_chain(KParams p) {
    __shared__ double s_acc[kChainGroups][6][kWave];
    if (batch_fatal(p)) return;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), grp = tid / kWave;
    const int j0 = p.colblk[blockIdx.x];
    const SysDesc& sd = p.sys[p.atom_sys[j0]];
    const int ab = sd.atom_begin;
    const int bj = (j0 - ab) / kWave;
    const int i = j0 + lane;  // pair slot
    const bool iv = i < sd.atom_end;
    const int pieces = (sd.nblk - bj + sd.pair_k - 1) / sd.pair_k;
    const long long T = p.colpart_total;
    double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (iv) {
        for (int k = grp; k < pieces; k += kChainGroups) {
            const double* in = p.pair_part + (size_t)k * 6 * p.N;
            a[0] += in[i];
            a[1] += in[p.N + i];
            a[2] += in[2 * p.N + i];
            a[3] += in[3 * p.N + i];
            a[4] += in[4 * p.N + i];
            if (ADJ) a[5] += in[5 * p.N + i];
        }
        for (int I = grp; I < bj; I += kChainGroups) {
            const long long o = colpart_index(sd, I, i - ab);
            a[0] += p.colpart[o];
            a[1] += p.colpart[T + o];
            a[2] += p.colpart[2 * T + o];
            a[3] += p.colpart[3 * T + o];
            if (ADJ) a[5] += p.colpart[4 * T + o];
        }
    }
    for (int v = 0; v < 6; ++v) s_acc[grp][v][lane] = a[v];
    __syncthreads();
    if (grp != 0 || !iv) return;
    const int ia = p.pair_atom[i];  // the atom of pair slot i
    double r[6];
    for (int v = 0; v < 6; ++v) {
        r[v] = s_acc[0][v][lane];
        for (int g = 1; g < kChainGroups; ++g) r[v] += s_acc[g][v][lane];
    }
    const double fx = r[0], fy = r[1], fz = r[2];
    double dedb = r[3], e = r[4];
    const double dq = r[5];
    const double q = p.charge[ia], B = p.born[ia], P = sd.prefactor;
    e += 0.5 * P * q * q / B;
    dedb += -0.5 * P * q * q / (B * B);
    p.atom_energy[ia] = e;
    if (!(fabs(fx) < kFixLimit && fabs(fy) < kFixLimit && fabs(fz) < kFixLimit))
        raise_flag(p, p.atom_sys[ia], FLAG_NONFINITE_RESULT);
    p.force_fix[ia] = (long long)llrint(fx * kFix32);
    p.force_fix[p.N + ia] = (long long)llrint(fy * kFix32);
    p.force_fix[2 * p.N + ia] = (long long)llrint(fz * kFix32);
    p.dedi[ia] = dedb * p.dbdi[ia];
    p.dedi[p.N + ia] = dedb * p.dbdi[p.N + ia];
    if (ADJ) {
        p.dedq[ia] = dq + P * q / B;
        p.dedr_fix[ia] = 0;
        // B = slope / D + shift, D = a0 (1/eta - I4) + a1 T^(1/4), T = 1/(4 eta^4) - I7
        const double I4 = p.integrals[sd.i4 * p.N + ia], I7 = p.integrals[sd.i7 * p.N + ia];
        const double eta = sd.eta, eta4 = eta * eta * eta * eta;
        const double T = 1.0 / (4.0 * eta4) - I7;
        const double T4 = pow(T, 0.25);
        const double D = sd.alpha0 * (1.0 / eta - I4) + sd.alpha1 * T4;
        const double g = -sd.slope / (D * D) * dedb;  // dE/dD
        const double dDdeta = -sd.alpha0 / (eta * eta) - sd.alpha1 * 0.25 * pow(T, -0.75) / (eta4 * eta);
        double* th = p.theta_atom;
        th[TH_PREFACTOR * p.N + ia] = P != 0.0 ? e / P : 0.0;
        th[TH_ALPHA0 * p.N + ia] = g * (1.0 / eta - I4);
        th[TH_ALPHA1 * p.N + ia] = g * T4;
        th[TH_ETA * p.N + ia] = g * dDdeta;
        th[TH_SLOPE * p.N + ia] = dedb / D;
        th[TH_SHIFT * p.N + ia] = dedb;
    }
}

// Energy of each system: one workgroup per system, fixed-order FP64 sum.
__global__ void __launch_bounds__(1024) k_energy(KParams p) {
    __shared__ double s_w[16];
    const int s = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    if (batch_fatal(p)) {
        if (tid == 0) p.energy[s] = 0.0;
        return;
    }
    const int ab = p.sys[s].atom_begin, ae = p.sys[s].atom_end;
    double acc = 0.0;
    for (int i = ab + tid; i < ae; i += blockDim.x) acc += p.atom_energy[i];
    acc = wave_sum(acc);
    if (lane == 0) s_w[wid] = acc;
    __syncthreads();
    if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += s_w[w];
        p.energy[s] = t;
        if (!isfinite(t)) raise_flag(p, s, FLAG_NONFINITE_RESULT);
    }
}

// ---------------------------------------------------------------------------
// Chain-rule force pass: one wave per cell.  Lanes form an 8 x 8 tile: lane = atom group
// (lane & 7) x point group (lane >> 3); a lane owns 8 atoms of the chunk (registers) and
// 8 points of each 64-point chunk (LDS).  For a point p of owner i with
// fac = sum_k dE/dI_k(i) w_q / r_q^(n_k - 2) and s = +1 (GBSW) / -1 (GBMV):
//   F_j     += s fac dV_p/dx_j                 (j != i)
//   F_owner -= s fac sum_j dV_p/dx_j           (translation invariance; written per point to vrec)
// (GBSWIntegral.cpp:271-314, GBMVIntegralTypeI.cpp:195-224, GBMVIntegralTypeII.cpp:221-282;
//  CharmmReferenceGBMV.cpp:376-383).  Atom partials are quantised to 2^-24 and summed
// exactly in FP64, so the per-cell totals do not depend on the point order.
// ADJ: also s fac dV_p/dR_j (GBSW: dR'_j) for every list atom including the owner, and
// the per-system parameter sums (VolAcc).
// ---------------------------------------------------------------------------

// Sum over the 8 lanes 8g..8g+7 (DPP: quad xor 1, quad xor 2, half-row mirror).
__device__ inline float sum8(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    return v;
}

// TRUNC (GBSW): the standalone force with a cutoff applies dB_i/dx_j only to list atoms j
// within the cutoff of the owner i (minimum image); the owner still takes the whole sum
// (CharmmReferenceGBSW.cpp:285-296).  The truncation shapes the forces only: with ADJ the
// radius and constant derivatives, which differentiate the energy, take every list atom.
template <int VK, bool ADJ, bool TRUNC>
__global__ void __launch_bounds__(64, (ADJ || TRUNC) ? 2 : 3) k_force(KParams p) {
    __shared__ float q_x[kWave], q_y[kWave], q_z[kWave];
    __shared__ float q_w[TRUNC ? 3 : 1][kWave];  // TRUNC: owner position relative to the cell corner
    __shared__ int q_tf[TRUNC ? kWave : 1];      // TRUNC: 0 every shell atom within c of the owner, 1 test, 2 none
    __shared__ float q_0[kWave], q_1[kWave], q_2[kWave], q_3[kWave], q_4[kWave], q_5[kWave];
    __shared__ int q_own[kWave], q_pid[kWave];
    // one axis of the per-lane atom partials at a time; rows padded by 8 doubles so the 8 rows
    // that one read instruction touches start 64 bytes apart (unpadded, 512-byte rows put 8
    // lanes on each bank)
    __shared__ double s_red[8][kWave + 8];
    if (batch_fatal(p)) return;
    const int lane = threadIdx.x, ag = lane & 7, pg = lane >> 3;
    const int nwork = p.work[VK];
    const int* worklist = p.worklist + VK * p.cell_total;
    const int* worklist_sys = p.worklist_sys + VK * p.cell_total;
    for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
        const int c = worklist[w];
        const int s = worklist_sys[w];
        const int nact = p.cell_act_count[c];
        if (nact == 0) continue;
        const SysDesc sd = p.sys[s];  // a copy: read once per cell
        const int na = p.cell_count_eff[c];
        const int ps = p.frec_base[c];
        const int as = p.cell_start[c];
        const int pt_base = sd.pt_base, abeg = sd.atom_begin, Q = sd.Q, qb = sd.quad_base;
        const float k1 = sd.k1, k3 = sd.k3, fclamp = sd.fclamp, sign_chain = sd.sign_chain;
        const float inv_sw = sd.inv_sw, P1x10 = 10.f * sd.P1;
        const bool trunc = TRUNC && sd.chain_trunc;
        BoxF bx;
        if (TRUNC) bx = box_f(p, s, sd);
        double pacc0 = 0.0, pacc1 = 0.0;  // ADJ: VolAcc partials of this lane
        for (int ab = 0; ab < na; ab += kWave) {
            float ax[8], ay[8], az[8], a0[8], a1[8], a2[8], aR[8];
            int aj[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int sl = ab + ag + 8 * k;
                float4 e = make_float4(0.f, 0.f, 0.f, __int_as_float(-1)), q = make_float4(0.f, 0.f, 0.f, 0.f);
                if (sl < na) {
                    e = p.ent_pos[as + sl];
                    q = p.ent_par[as + sl];
                }
                ax[k] = e.x;
                ay[k] = e.y;
                az[k] = e.z;
                aj[k] = __float_as_int(e.w);
                a0[k] = q.x;
                a1[k] = q.y;
                a2[k] = q.z;
                aR[k] = q.w;
            }
            double acc[8][3];
            double accr[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                acc[k][0] = acc[k][1] = acc[k][2] = 0.0;
                accr[k] = 0.0;
            }
            // the atoms' registers complete here: otherwise every (point row, atom) block waits
            // for them, and that wait (vmcnt counts stores too) also waits out the previous
            // row's owner-force store
            __builtin_amdgcn_s_waitcnt(kWaitVmem);
            for (int pb = 0; pb < nact; pb += kWave) {
                const int cntp = min(kWave, nact - pb);
                __syncthreads();
                if (lane < cntp) {
                    const int o = ps + pb + lane;
                    const float4 r0 = p.frec0[o];
                    const float4 r1 = p.frec1[o];
                    const uint32_t pk = __float_as_uint(r0.w);
                    const int i = (int)(pk >> kQBits), q = (int)(pk & ((1u << kQBits) - 1));
                    const double fac = p.dedi[i] * p.quad_pw[qb + q] + p.dedi[p.N + i] * p.quad_pw[p.quad_rows + qb + q];
                    const float facs = sign_chain * (float)fac;
                    q_x[lane] = r0.x;
                    q_y[lane] = r0.y;
                    q_z[lane] = r0.z;
                    if (TRUNC) {
                        // shell atoms lie within max(R' + w) of the point, the point r_q from its
                        // owner: r_q + reach below c keeps them all, r_q - reach above c none
                        // (not periodic: another image may be nearer)
                        const int row = qb + q;
                        q_w[0][lane] = r0.x - (float)p.quad_off[3 * row];
                        q_w[1][lane] = r0.y - (float)p.quad_off[3 * row + 1];
                        q_w[2][lane] = r0.z - (float)p.quad_off[3 * row + 2];
                        const float rq = fast_sqrt(p.quad_r2[row]);
                        q_tf[lane] = !trunc || rq < sd.trunc_in ? 0 : (rq > sd.trunc_out ? 2 : 1);
                    }
                    q_own[lane] = i;
                    q_pid[lane] = pt_base + (i - abeg) * Q + q;
                    if (VK == VOL_GBSW) {
                        q_0[lane] = facs;
                        q_1[lane] = r1.x;
                    } else if (VK == VOL_GBMV2) {
                        const float4 r2 = p.frec2[o];
                        q_0[lane] = facs * r1.x;
                        q_1[lane] = facs * r1.y;
                        q_2[lane] = facs * r1.z;
                        q_3[lane] = r1.w;
                        q_4[lane] = r2.x;
                        q_5[lane] = r2.y;
                    } else {
                        q_0[lane] = facs * r1.x;
                    }
                }
                __syncthreads();
                const int tmax = (cntp + 7) >> 3;  // point slots pg + 8 t < cntp exist
                for (int t = 0; t < tmax; ++t) {
                    const int m = pg + 8 * t;
                    float ox = 0.f, oy = 0.f, oz = 0.f;
                    if (m < cntp) {
                        const float px = q_x[m], py = q_y[m], pz = q_z[m];
                        const int own = q_own[m];
                        const float d0 = q_0[m], d1 = q_1[m];
                        const float d2c = VK == VOL_GBMV2 ? q_2[m] : 0.f;
                        const float vx = VK == VOL_GBMV2 ? q_3[m] : 0.f;
                        const float vy = VK == VOL_GBMV2 ? q_4[m] : 0.f;
                        const float vz = VK == VOL_GBMV2 ? q_5[m] : 0.f;
                        const float wx = TRUNC ? q_w[0][m] : 0.f, wy = TRUNC ? q_w[1][m] : 0.f,
                                    wz = TRUNC ? q_w[2][m] : 0.f;
                        const int tf = TRUNC ? q_tf[m] : 0;
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            if (aj[k] < 0) continue;
                            const bool is_own = aj[k] == own;
                            if (!ADJ && is_own) continue;
                            const float dx = ax[k] - px, dy = ay[k] - py, dz = az[k] - pz;
                            const float dd = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                            float sv = 0.f, c3F = 0.f;
                            if (VK == VOL_GBSW) {
                                // shell atoms only: (R'-w)^2 < d^2 < (R'+w)^2
                                if (dd < a2[k] && dd > a1[k]) {
                                    const float d = fast_sqrt(dd);
                                    const float x = d - a0[k];
                                    const float h = fmaf(x, fmaf(-k3 * x, x, k1), 0.5f);
                                    if (h > 0.f) {
                                        const float dh = fmaf(-3.f * k3 * x, x, k1);
                                        const float gV = d0 * (d1 * fast_rcp(h));  // s fac V / H_j
                                        sv = gV * dh * fast_rcp(d);
                                        if (ADJ) {
                                            // dV/dR'_j = -V H'_j/H_j; dV/dw = V (dH_j/dw)/H_j
                                            accr[k] += quant(-gV * dh);
                                            const float xw = x * inv_sw;
                                            pacc0 += quant(gV * 0.75f * xw * inv_sw * fmaf(xw, xw, -1.f));
                                        }
                                    }
                                }
                            } else if (VK == VOL_GBMV2) {
                                const float tt = fmaf(100.f, dd, a1[k]);
                                const float rt = fast_rcp(tt);
                                const float ct = a0[k] * rt;
                                const float F = fminf(ct * ct, fclamp);
                                const float D = -a2[k] * rt * rt * rt;
                                const float vdr = fmaf(vx, dx, fmaf(vy, dy, vz * dz));
                                sv = fmaf(d0, D, fmaf(d1 * F, fmaf(D, dd, F), d2c * vdr * D));
                                c3F = d2c * F;
                                if (ADJ) {
                                    // s fac dV/dF_j; F = (C/t)^2, t = b + 100 d^2, b = C - 100 R^2
                                    const float G = fmaf(d1 * dd, F, fmaf(d2c, vdr, d0));
                                    const float dFdC = 2.f * ct * rt, dFdb = -2.f * ct * ct * rt;
                                    const float gab = G * (dFdC + dFdb);
                                    accr[k] += quant(fmaf(G * dFdb, -200.f * aR[k], gab * P1x10));
                                    pacc0 += quant(gab);
                                    pacc1 += quant(gab * 10.f * aR[k]);
                                }
                            } else {
                                const float gam = a0[k];
                                const float ex = __expf(gam * (dd * dd));
                                sv = d0 * gam * ex * 4.f * dd;
                                if (ADJ) {
                                    // dS/dgamma_j = d^4 e_j; dgamma_j/dR_j = -4 gamma_j/R_j
                                    const float tg = d0 * (dd * dd) * ex;
                                    accr[k] += quant(tg * (-4.f * gam / aR[k]));
                                    pacc0 += quant(tg * gam);
                                }
                            }
                            if (ADJ && is_own) continue;  // the owner moves with its points
                            const float cx = fmaf(sv, dx, c3F * vx);
                            const float cy = fmaf(sv, dy, c3F * vy);
                            const float cz = fmaf(sv, dz, c3F * vz);
                            bool keep = !TRUNC || tf != 2;
                            if (TRUNC && tf == 1 && sv != 0.f) {
                                float ex = ax[k] - wx, ey = ay[k] - wy, ez = az[k] - wz;
                                if (sd.periodic) {
                                    ex = min_image(ex, bx.L[0], bx.inv[0]);
                                    ey = min_image(ey, bx.L[1], bx.inv[1]);
                                    ez = min_image(ez, bx.L[2], bx.inv[2]);
                                }
                                keep = within_cutoff(p, s, sd, fmaf(ex, ex, fmaf(ey, ey, ez * ez)), own, aj[k]);
                            }
                            if (keep) {
                                acc[k][0] += quant(cx);
                                acc[k][1] += quant(cy);
                                acc[k][2] += quant(cz);
                            }
                            ox += cx;
                            oy += cy;
                            oz += cz;
                        }
                    }
                    {
                        ox = sum8(ox);
                        oy = sum8(oy);
                        oz = sum8(oz);
                        if (ag == 0 && m < cntp) {
                            const int pid = q_pid[m];
                            if (ab == 0) {
                                p.vrec[pid] = make_float4(-ox, -oy, -oz, 0.f);
                            } else {
                                float4 v = p.vrec[pid];
                                p.vrec[pid] = make_float4(v.x - ox, v.y - oy, v.z - oz, 0.f);
                            }
                        }
                    }
                }
            }
            // exact sums over the point groups: list atom ab + L (atom group L & 7, slot L >> 3)
            // is summed by lane L over the lanes (L & 7) + 8 pg, one axis at a time through LDS;
            // then every lane adds its atom's totals (the partials are integers on the 2^-24
            // grid, so any order gives the same sum)
            const int my_sl = ab + lane;
            const int my_atom = my_sl < na ? __float_as_int(p.ent_pos[as + my_sl].w) : -1;
            for (int d = 0; d < (ADJ ? 4 : 3); ++d) {
#pragma unroll
                for (int k = 0; k < 8; ++k) s_red[k][lane] = d < 3 ? acc[k][d] : accr[k];
                __syncthreads();
                double tot = 0.0;
#pragma unroll
                for (int g = 0; g < 8; ++g) tot += s_red[lane >> 3][(lane & 7) + 8 * g];
                __syncthreads();
                if (my_atom >= 0 && tot != 0.0) {
                    if (!(fabs(tot) < kFixLimit24))  // NaN, inf, or beyond the int64 range
                        raise_flag(p, s, FLAG_NONFINITE_RESULT);
                    else if (d < 3)
                        add_fix(&p.force_fix[d * p.N + my_atom], tot);
                    else
                        add_fix(&p.dedr_fix[my_atom], tot);
                }
            }
        }
        if (ADJ) {
            pacc0 = wave_sum(pacc0);
            pacc1 = wave_sum(pacc1);
            if (lane == 0) {
                long long* va = p.vacc_fix + ((long long)s * kVaccCopies + (w & (kVaccCopies - 1))) * kVaccStride;
                const int i0 = VK == VOL_GBSW ? VA_W : VK == VOL_GBMV2 ? VA_A : VA_GAMMA;
                if (pacc0 != 0.0) add_fix(&va[i0], pacc0);
                if (VK == VOL_GBMV2 && pacc1 != 0.0) add_fix(&va[VA_AR], pacc1);
            }
        }
    }
}

// Owner forces: one wave per atom, fixed-order FP64 sum of its points' owner forces, read
// only where the sign bit of V marks a point with a chain-rule factor (V >= 0 always).
// ADJ (GBMV): the sigmoid parts of dE/dbeta, dE/dlambda, dE/dS0 with
// V = 1/(1 + exp(beta (S - lambda))): dV/dbeta = -V(1-V)(S - lambda), dV/dlambda = V(1-V) beta,
// dV/dS0 = -V(1-V) beta S/S0; dE/dtheta = -sum s fac dV/dtheta.
template <bool ADJ>
__global__ void __launch_bounds__(64) k_owner_reduce(KParams p) {
    if (batch_fatal(p)) return;
    const int i = blockIdx.x, lane = threadIdx.x;
    const int s = p.atom_sys[i];
    const SysDesc& sd = p.sys[s];
    const int Q = sd.Q, qb = sd.quad_base;
    const int pbase = sd.pt_base + (i - sd.atom_begin) * Q;
    const bool gbsw = sd.vol_kind == VOL_GBSW;
    const double e4 = p.dedi[i], e7 = p.dedi[p.N + i];
    const double sc = sd.sign_chain;
    const double beta = sd.beta, lambda = sd.lambda;
    double a[3] = {0.0, 0.0, 0.0};
    double tb = 0.0, tl = 0.0, ts = 0.0;
    auto point = [&](int q, float vv) {
        const int pid = pbase + q;
        const bool marked = __float_as_int(vv) < 0;
        const float4 v = p.vrec[marked ? pid : pbase];
        a[0] += marked ? (double)v.x : 0.0;
        a[1] += marked ? (double)v.y : 0.0;
        a[2] += marked ? (double)v.z : 0.0;
        if (ADJ && !gbsw) {
            const float S = p.vsig[pid];
            if (!isnan(S)) {
                const double facs = sc * (e4 * p.quad_pw[qb + q] + e7 * p.quad_pw[p.quad_rows + qb + q]);
                const double V = volume_of(vv);
                const double u = V * (1.0 - V);
                tb += facs * u * ((double)S - lambda);
                tl -= facs * u * beta;
                ts += facs * u * beta * (double)S;
            }
        }
    };
    if ((Q & 3) == 0) {  // 4 points per lane and load, as k_integrals
        const float4* v4 = reinterpret_cast<const float4*>(p.vval + pbase);
        for (int q4 = lane; q4 < (Q >> 2); q4 += kWave) {
            const float4 w = v4[q4];
            point(4 * q4, w.x);
            point(4 * q4 + 1, w.y);
            point(4 * q4 + 2, w.z);
            point(4 * q4 + 3, w.w);
        }
    } else {
#pragma unroll 4
        for (int q = lane; q < Q; q += kWave) point(q, p.vval[pbase + q]);
    }
    for (int d = 0; d < 3; ++d) a[d] = wave_sum(a[d]);
    if (ADJ) {
        tb = wave_sum(tb);
        tl = wave_sum(tl);
        ts = wave_sum(ts);
    }
    if (lane == 0) {
        if (!(fabs(a[0]) < kFixLimit && fabs(a[1]) < kFixLimit && fabs(a[2]) < kFixLimit))
            raise_flag(p, s, FLAG_NONFINITE_RESULT);
        for (int d = 0; d < 3; ++d) p.force_fix[d * p.N + i] += (long long)llrint(a[d] * kFix32);
        if (ADJ) {
            p.theta_atom[TH_BETA * p.N + i] = tb;
            p.theta_atom[TH_LAMBDA * p.N + i] = tl;
            p.theta_atom[TH_S0 * p.N + i] = sd.S0 != 0.f ? ts / (double)sd.S0 : 0.0;
        }
    }
}

// Adjoint, per atom: dE/dR from the fixed-point chain sums; GBSW also the per-atom parts of
// dE/drscale and dE/droffset (R' = rscale (R + roffset)).
__global__ void k_adjoint_atoms(KParams p) {
    if (batch_fatal(p)) return;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= p.N) return;
    const SysDesc& sd = p.sys[p.atom_sys[i]];
    const double g = -(double)p.dedr_fix[i] / kFix32;  // dE/dR'_j (GBSW) or dE/dR_j
    double* th = p.theta_atom;
    if (sd.vol_kind == VOL_GBSW) {
        p.dedr[i] = g * (double)sd.rscale;
        th[TH_RSCALE * p.N + i] = g * (p.radius[i] + (double)sd.roffset);
        th[TH_ROFFSET * p.N + i] = g * (double)sd.rscale;
    } else {
        p.dedr[i] = g;
        th[TH_RSCALE * p.N + i] = 0.0;
        th[TH_ROFFSET * p.N + i] = 0.0;
    }
}

// Adjoint, per system: fixed-order sums of the per-atom parts plus the force-pass sums.  One
// wave per sum (11 per-atom rows, then the VA_COUNT accumulators, lane c reading copy c), all
// in flight together; thread 0 combines them.
constexpr int kThetaRows = 11;
constexpr int kThetaThreads = 1024;
static_assert(kVaccCopies == kWave, "one lane per accumulator copy");
static_assert((kThetaRows + VA_COUNT) * kWave <= kThetaThreads, "one wave per sum");

__global__ void __launch_bounds__(kThetaThreads) k_theta_reduce(KParams p) {
    __shared__ double s_row[kThetaRows];
    __shared__ long long s_va[VA_COUNT];
    const int s = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const SysDesc& sd = p.sys[s];
    if (batch_fatal(p)) {
        if (tid < TH_COUNT) p.theta[s * TH_COUNT + tid] = 0.0;
        return;
    }
    const int atom_rows[kThetaRows] = {TH_PREFACTOR, TH_ALPHA0, TH_ALPHA1, TH_ETA,   TH_SLOPE, TH_SHIFT,
                                       TH_RSCALE,    TH_ROFFSET, TH_BETA,  TH_LAMBDA, TH_S0};
    if (wid < kThetaRows) {
        const double* row = p.theta_atom + (size_t)atom_rows[wid] * p.N;
        double acc[4] = {0.0, 0.0, 0.0, 0.0};  // four loads in flight per lane
        int i = sd.atom_begin + lane;
        for (; i + 3 * kWave < sd.atom_end; i += 4 * kWave)
            for (int u = 0; u < 4; ++u) acc[u] += row[i + u * kWave];
        for (int u = 0; i < sd.atom_end; i += kWave, ++u) acc[u] += row[i];
        const double sum = wave_sum((acc[0] + acc[1]) + (acc[2] + acc[3]));
        if (lane == 0) s_row[wid] = sum;
    } else if (wid < kThetaRows + VA_COUNT) {
        const int k = wid - kThetaRows;
        long long v = p.vacc_fix[((long long)s * kVaccCopies + lane) * kVaccStride + k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
        if (lane == 0) s_va[k] = v;
    }
    __syncthreads();
    if (tid == 0) {
        double* th = p.theta + s * TH_COUNT;
        for (int r = 0; r < kThetaRows; ++r) th[atom_rows[r]] = s_row[r];
        th[TH_SW] = th[TH_GAMMA0] = th[TH_P1] = th[TH_P2] = 0.0;
        if (sd.vol_kind == VOL_GBSW) {
            th[TH_SW] = -(double)s_va[VA_W] / kFix32;
        } else if (sd.vol_kind == VOL_GBMV2) {
            th[TH_P1] = -(double)s_va[VA_AR] / kFix32;
            th[TH_P2] = -(double)s_va[VA_A] / kFix32;
        } else {
            const double gsum = -(double)s_va[VA_GAMMA] / kFix32;
            th[TH_GAMMA0] = gsum / (double)sd.gamma0;
            th[TH_LAMBDA] += gsum / ((double)sd.lambda * log((double)sd.lambda));
        }
    }
}

// Points, point/atom pairs of the volume pass, and the active points and point/atom pairs
// of the chain-rule pass (counts of the last evaluation; diagnostics only).
__global__ void k_stats(KParams p, double* out) {
    if (batch_fatal(p)) return;
    double np = 0.0, pairs = 0.0, nact = 0.0, fpairs = 0.0;
    const int stride = gridDim.x * blockDim.x;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < p.cell_total; c += stride) {
        const double a = p.cell_pt_count[c];
        const double act = a > 0.0 ? p.cell_act_count[c] : 0.0;  // written only for cells with points
        np += a;
        pairs += a * p.cell_count_eff[c];
        nact += act;
        fpairs += act * p.cell_count_eff[c];
    }
    np = wave_sum(np);
    pairs = wave_sum(pairs);
    nact = wave_sum(nact);
    fpairs = wave_sum(fpairs);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        atomicAdd(&out[0], np);
        atomicAdd(&out[1], pairs);
        atomicAdd(&out[4], nact);
        atomicAdd(&out[5], fpairs);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double nc = 0.0;
        for (int s = 0; s < p.S; ++s) nc += p.geom[s].ncell;
        out[2] = nc;
        out[3] = p.cell_start[p.cell_total];
    }
}

template <bool ADJ>
void launch_force(const KParams& p, int vk, int blocks, hipStream_t stream) {
    switch (vk) {
    case VOL_GBSW:
        if (p.any_trunc)
            hipLaunchKernelGGL((k_force<VOL_GBSW, ADJ, true>), dim3(blocks), dim3(64), 0, stream, p);
        else
            hipLaunchKernelGGL((k_force<VOL_GBSW, ADJ, false>), dim3(blocks), dim3(64), 0, stream, p);
        break;
    case VOL_GBMV1: hipLaunchKernelGGL((k_force<VOL_GBMV1, ADJ, false>), dim3(blocks), dim3(64), 0, stream, p); break;
    default: hipLaunchKernelGGL((k_force<VOL_GBMV2, ADJ, false>), dim3(blocks), dim3(64), 0, stream, p); break;
    }
}

}  // namespace

hipError_t launch_evaluate(const KParams& p, const bool* kinds, hipStream_t stream, const hipEvent_t* marks) {
    hipError_t err;
    auto mark = [&](int k) {
        if (marks) (void)hipEventRecord(marks[k], stream);
    };
    const bool adj = p.adjoint != 0;
    mark(0);
    const int clear_blocks = std::max(1, std::min((p.cell_total + 255) / 256, 4096));
    hipLaunchKernelGGL(k_clear, dim3(clear_blocks), dim3(256), 0, stream, p);
    hipLaunchKernelGGL(k_bbox, dim3(p.bbox_slices, p.S), dim3(256), 0, stream, p);
    hipLaunchKernelGGL(k_grid_fill, dim3(p.N), dim3(256), 0, stream, p);
    err = scan_exclusive(p.cell_fill, p.cell_start, p.cell_total, p.scan_tmp, p.entry_cap, p.flags_all,
                         FLAG_ENTRY_OVERFLOW, stream);
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(k_grid_sort, dim3(std::max(1, std::min(p.cell_total, 65536))), dim3(64), 0, stream, p);
    mark(1);
    if (p.any_cells) hipLaunchKernelGGL(k_point_bin<false>, dim3(p.N), dim3(bin_threads<false>()), 0, stream, p);
    if (p.any_bricks) hipLaunchKernelGGL(k_point_bin<true>, dim3(p.N), dim3(bin_threads<true>()), 0, stream, p);
    if (p.any_bricks)
        hipLaunchKernelGGL(k_point_split, dim3(std::max(1, std::min(p.brick_max, 4096)), p.S, kSplitParts), dim3(256),
                           0, stream, p);
    // records of the active points of cell c start at the exclusive prefix sum of the point counts
    // (the counters gathered into cell_pt_count); k_worklist adds the tile offsets
    {
        const int nblocks = (p.cell_total + kScanTile - 1) / kScanTile;
        if (nblocks > kScanTile || nblocks < 1) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_scan_tiles, dim3(nblocks), dim3(kScanThreads), 0, stream, p.cell_ctr, p.frec_base,
                           p.cell_total, p.scan_tmp, p.ctr_perm, p.cell_pt_count);
    }
    hipLaunchKernelGGL(k_worklist, dim3(std::max(1, std::min((p.cell_total + kWorkThreads - 1) / kWorkThreads, 4096))),
                       dim3(kWorkThreads), 0, stream, p);
    const int pb = std::max(1, std::min(p.cell_total, p.persist_blocks));
    mark(2);
    if (kinds[VOL_GBSW]) hipLaunchKernelGGL(k_volume<VOL_GBSW>, dim3(pb), dim3(64), 0, stream, p);
    if (kinds[VOL_GBMV1]) hipLaunchKernelGGL(k_volume<VOL_GBMV1>, dim3(pb), dim3(64), 0, stream, p);
    if (kinds[VOL_GBMV2]) hipLaunchKernelGGL(k_volume<VOL_GBMV2>, dim3(pb), dim3(64), 0, stream, p);
    mark(3);
    hipLaunchKernelGGL(k_integrals, dim3(p.N), dim3(64), 0, stream, p);
    mark(4);
    if (p.custom) {
        // general expressions (one system): values, pair tiles, chain
        const CustomLaunch& c = *p.custom;
        int N = c.N, nb = c.nb, method = c.method;
        double cut2 = c.cut2;
        const double* pos = p.pos;
        const double* integ = p.integrals;
        const double *prm = c.prm, *glob = c.glob, *box = c.box;
        double *vals = c.vals, *dvdi = c.dvdi, *esp = c.esp, *despdi = c.despdi, *part = c.part;
        double *dvdth = c.dvdth, *despdth = c.despdth, *pdth = c.pdth, *dvdx = c.dvdx, *despdx = c.despdx;
        const int *ex_ptr = c.ex_ptr, *ex_idx = c.ex_idx;
        double* atom_energy = p.atom_energy;
        long long* force_fix = p.force_fix;
        double* dedi = p.dedi;
        void* va[] = {&N, &prm, &integ, &glob, &vals, &dvdi, &esp, &despdi, &dvdth, &despdth, &pos, &dvdx, &despdx};
        err = hipModuleLaunchKernel(c.values, (N + 255) / 256, 1, 1, 256, 1, 1, 0, stream, va, nullptr);
        if (err != hipSuccess) return err;
        void* pa[] = {&N, &nb, &pos, &prm, &vals, &glob, &method, &cut2, &box, &ex_ptr, &ex_idx, &part};
        err = hipModuleLaunchKernel(c.pairs, nb, nb, 1, c.pair_threads, 1, 1, 0, stream, pa, nullptr);
        if (err != hipSuccess) return err;
        mark(5);
        void* ca[] = {&N,    &nb,        &part, &dvdi,    &esp,    &despdi, &atom_energy,
                      &force_fix, &dedi, &dvdth, &despdth, &pdth, &dvdx, &despdx};
        err = hipModuleLaunchKernel(c.chain, nb, 1, 1, c.chain_threads, 1, 1, 0, stream, ca, nullptr);
        if (err != hipSuccess) return err;
    } else if (adj) {
        if (p.any_cut || p.any_excl) {
            hipLaunchKernelGGL(k_block_box, dim3(p.n_colblocks), dim3(64), 0, stream, p);
            hipLaunchKernelGGL((k_pairs<true, true>), dim3(p.n_tiles), dim3(64), 0, stream, p);
        } else {
            hipLaunchKernelGGL((k_pairs<true, false>), dim3(p.n_tiles), dim3(64), 0, stream, p);
        }
    } else if (p.any_cut || p.any_excl) {
        hipLaunchKernelGGL(k_block_box, dim3(p.n_colblocks), dim3(64), 0, stream, p);
        hipLaunchKernelGGL((k_pairs<false, true>), dim3(p.n_tiles), dim3(64), 0, stream, p);
    } else {
        hipLaunchKernelGGL((k_pairs<false, false>), dim3(p.n_tiles), dim3(64), 0, stream, p);
    }
    if (!p.custom) {
        mark(5);
        if (adj)
            hipLaunchKernelGGL(k_atom_chain<true>, dim3(p.n_colblocks), dim3(kChainGroups * kWave), 0, stream, p);
        else
            hipLaunchKernelGGL(k_atom_chain<false>, dim3(p.n_colblocks), dim3(kChainGroups * kWave), 0, stream, p);
    }
    hipLaunchKernelGGL(k_energy, dim3(p.S), dim3(1024), 0, stream, p);
    mark(6);
    for (int vk = 0; vk < VOL_KINDS; ++vk) {
        if (!kinds[vk]) continue;
        if (adj)
            launch_force<true>(p, vk, pb, stream);
        else
            launch_force<false>(p, vk, pb, stream);
    }
    mark(7);
    if (adj) {
        hipLaunchKernelGGL(k_owner_reduce<true>, dim3(p.N), dim3(64), 0, stream, p);
        hipLaunchKernelGGL(k_adjoint_atoms, dim3((p.N + 255) / 256), dim3(256), 0, stream, p);
        hipLaunchKernelGGL(k_theta_reduce, dim3(p.S), dim3(kThetaThreads), 0, stream, p);
    } else {
        hipLaunchKernelGGL(k_owner_reduce<false>, dim3(p.N), dim3(64), 0, stream, p);
    }
    mark(8);
    return hipGetLastError();
}

namespace {

// Positions from an OpenMM posq buffer (float4 or double4 per slot) in the platform's order.
__global__ void k_gather_posq(double* pos, const void* posq, int posq_double, const int* atom_slot, int n) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int slot = atom_slot ? atom_slot[a] : a;
    if (posq_double) {
        const double4 v = static_cast<const double4*>(posq)[slot];
        pos[3 * a] = v.x;
        pos[3 * a + 1] = v.y;
        pos[3 * a + 2] = v.z;
    } else {
        const float4 v = static_cast<const float4*>(posq)[slot];
        pos[3 * a] = (double)v.x;
        pos[3 * a + 1] = (double)v.y;
        pos[3 * a + 2] = (double)v.z;
    }
}

// Forces added to an OpenMM fixed-point force buffer (axis-major, padded_n per axis); one
// thread per atom, so each slot has one writer.
__global__ void k_add_forces_fixed(long long* dst, int padded_n, const long long* force_fix, const int* atom_slot,
                                   int n) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int slot = atom_slot ? atom_slot[a] : a;
    for (int d = 0; d < 3; ++d) dst[(long long)d * padded_n + slot] += force_fix[(long long)d * n + a];
}

}  // namespace

hipError_t launch_gather_posq(double* pos, const void* posq, int posq_double, const int* atom_slot, int n,
                              hipStream_t stream) {
    hipLaunchKernelGGL(k_gather_posq, dim3((n + 255) / 256), dim3(256), 0, stream, pos, posq, posq_double, atom_slot, n);
    return hipGetLastError();
}

hipError_t launch_add_forces_fixed(long long* dst, int padded_n, const long long* force_fix, const int* atom_slot,
                                   int n, hipStream_t stream) {
    hipLaunchKernelGGL(k_add_forces_fixed, dim3((n + 255) / 256), dim3(256), 0, stream, dst, padded_n, force_fix,
                       atom_slot, n);
    return hipGetLastError();
}

hipError_t launch_stats(const KParams& p, double* out6, hipStream_t stream) {
    hipLaunchKernelGGL(k_stats, dim3(256), dim3(256), 0, stream, p, out6);
    return hipGetLastError();
}

}  // namespace gbi
