/*
 * gb_oracle.c -- FP64 CPU restatement of the integral Generalized-Born path of
 * ymwangg/OpenMMCustomGBIntegralForce (Reference platform).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  It is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * library never links it.
 *
 * Two loop structures:
 *   dense = 1  the reference's own: q-outer / atom-inner integral loop with dense
 *              dI_k(i)/dx_j arrays (GBMV form, CharmmReferenceGBMV.cpp:48,376-383) or
 *              dense dB_i/dx_j (standalone GBSW, CharmmReferenceGBSW.cpp:23,285-296).
 *              This is the CPU baseline ("port" of the Reference platform).
 *   dense = 0  sparse: integrals per atom, then the chain rule recomputed per
 *              quadrature point once dE/dI is known (memory O(N Q)); optional OpenMP.
 *
 * Decisions where the reference is undefined or inconsistent (each one is tested):
 *   - grid cells outside [0, n-1] are never written (the reference indexes them, which
 *     is out of bounds for atoms with R > 0.03 nm on the bounding-box faces,
 *     CustomGBIntegral.cpp:198-227);
 *   - a point outside the grid has an empty list (CustomGBIntegral.cpp:314-317 leaves
 *     the list unset; GBMVIntegralTypeII.cpp:131 reads it uninitialised);
 *   - the standalone GBSW accumulators are zeroed every call
 *     (CharmmReferenceGBSW.cpp:235,562 never clear them);
 *   - GBMV-II brute force uses the lookup-path gradient (GBMVIntegralTypeII.cpp:243),
 *     not the one missing the x100 unit factor (GBMVIntegralTypeII.cpp:304);
 *   - GBMV-II with |v|^2 < vguard takes the S -> +inf limit (V = 1 for beta < 0, V = 0 for
 *     beta > 0; no gradient);
 *   - GBSW smoothing H <= 0 inside the switching shell is treated as burial (V = 0).
 */
#include "gb_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_ERR_ARG (-1)
#define ORC_ERR_MEM (-2)

/* ------------------------------------------------------------------------- */
/* Quadrature                                                                */
/* ------------------------------------------------------------------------- */

/* Legendre P_n(z) and P_n'(z) by the three-term recurrence. */
static void legendre_eval(int n, double z, double* pn, double* dpn) {
    double pkm1 = 1.0, pk = z;
    if (n == 0) { *pn = 1.0; *dpn = 0.0; return; }
    for (int k = 2; k <= n; ++k) {
        double pkp1 = ((2.0 * k - 1.0) * z * pk - (k - 1.0) * pkm1) / k;
        pkm1 = pk;
        pk = pkp1;
    }
    *pn = pk;
    *dpn = n * (z * pk - pkm1) / (z * z - 1.0);
}

/* Gauss-Legendre nodes (ascending) and weights on [x0, x1] by Newton iteration on
 * P_n from the Chebyshev-like first guess, as CharmmQuadrature.h:47-75. */
int orc_gauss_legendre(double x0, double x1, int n, double* x, double* w) {
    if (n <= 0 || !x || !w) return ORC_ERR_ARG;
    const double pi = 3.14159265358979323846;
    const double mid = 0.5 * (x1 + x0), half = 0.5 * (x1 - x0);
    for (int i = 1; i <= (n + 1) / 2; ++i) {
        double z = cos(pi * (i - 0.25) / (n + 0.5));
        double pn, dpn;
        for (int it = 0; it < 100; ++it) {
            legendre_eval(n, z, &pn, &dpn);
            double dz = pn / dpn;
            z -= dz;
            if (fabs(dz) <= 3.0e-14) break;
        }
        legendre_eval(n, z, &pn, &dpn);
        x[i - 1] = mid - half * z;
        x[n - i] = mid + half * z;
        w[i - 1] = 2.0 * half / ((1.0 - z * z) * dpn * dpn);
        w[n - i] = w[i - 1];
    }
    return 0;
}

/* One octahedral orbit of a Lebedev rule (the gen_oh codes of CharmmQuadrature.h:183-776).
 * code 1: (+-1,0,0)            6 pts     code 2: (0,+-a,+-a), a=1/sqrt2     12 pts
 * code 3: (+-a,+-a,+-a) a=1/sqrt3  8     code 4: (+-a,+-a,+-b), b=sqrt(1-2a^2) 24
 * code 5: (+-a,+-b,0), b=sqrt(1-a^2)  24 */
static int lebedev_orbit(int code, double a, double v, double* x, double* y, double* z, double* w) {
    int n = 0;
#define PUT(X, Y, Z) do { x[n] = (X); y[n] = (Y); z[n] = (Z); w[n] = v; ++n; } while (0)
    if (code == 1) {
        for (int s = -1; s <= 1; s += 2) { PUT(s, 0, 0); PUT(0, s, 0); PUT(0, 0, s); }
    } else if (code == 2) {
        double c = sqrt(0.5);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2) {
                PUT(0, s1 * c, s2 * c); PUT(s1 * c, 0, s2 * c); PUT(s1 * c, s2 * c, 0);
            }
    } else if (code == 3) {
        double c = sqrt(1.0 / 3.0);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2)
                for (int s3 = -1; s3 <= 1; s3 += 2) PUT(s1 * c, s2 * c, s3 * c);
    } else if (code == 4) {
        double b = sqrt(1.0 - 2.0 * a * a);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2)
                for (int s3 = -1; s3 <= 1; s3 += 2) {
                    PUT(s1 * a, s2 * a, s3 * b); PUT(s1 * a, s2 * b, s3 * a); PUT(s1 * b, s2 * a, s3 * a);
                }
    } else if (code == 5) {
        double b = sqrt(1.0 - a * a);
        for (int s1 = -1; s1 <= 1; s1 += 2)
            for (int s2 = -1; s2 <= 1; s2 += 2) {
                PUT(s1 * a, s2 * b, 0); PUT(s1 * b, s2 * a, 0);
                PUT(s1 * a, 0, s2 * b); PUT(s1 * b, 0, s2 * a);
                PUT(0, s1 * a, s2 * b); PUT(0, s1 * b, s2 * a);
            }
    } else {
        return -1;
    }
#undef PUT
    return n;
}

/* Lebedev-Laikov rules of 6..50 points (orbit weights as published, and as tabulated in
 * ld0006..ld0050, CharmmQuadrature.h:1000-1200).  Weights sum to 1. */
int orc_lebedev(int npts, double* x, double* y, double* z, double* w) {
    int n = 0;
    switch (npts) {
    case 6:
        n += lebedev_orbit(1, 0, 1.0 / 6.0, x + n, y + n, z + n, w + n);
        break;
    case 14:
        n += lebedev_orbit(1, 0, 0.6666666666666667e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(3, 0, 0.7500000000000000e-1, x + n, y + n, z + n, w + n);
        break;
    case 26:
        n += lebedev_orbit(1, 0, 0.4761904761904762e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(2, 0, 0.3809523809523810e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(3, 0, 0.3214285714285714e-1, x + n, y + n, z + n, w + n);
        break;
    case 38:
        n += lebedev_orbit(1, 0, 0.9523809523809524e-2, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(3, 0, 0.3214285714285714e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(5, 0.4597008433809831, 0.2857142857142857e-1, x + n, y + n, z + n, w + n);
        break;
    case 50:
        n += lebedev_orbit(1, 0, 0.1269841269841270e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(2, 0, 0.2257495590828924e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(3, 0, 0.2109375000000000e-1, x + n, y + n, z + n, w + n);
        n += lebedev_orbit(4, 0.3015113445777636, 0.2017333553791887e-1, x + n, y + n, z + n, w + n);
        break;
    default:
        return ORC_ERR_ARG;
    }
    return n == npts ? 0 : ORC_ERR_ARG;
}

/* order_table (CharmmQuadrature.h:6801-6862), first five rules. */
int orc_lebedev_rule_npts(int rule) {
    static const int table[5] = {6, 14, 26, 38, 50};
    if (rule < 1 || rule > 5) return ORC_ERR_ARG;
    return table[rule - 1];
}

/* Point table: idx = radial*nL + angular, (r u_x, r u_y, r u_z, r, w_r w_s)
 * (GBSWIntegral.cpp:41-56, CharmmReferenceGBSW.cpp:31-46). */
int orc_quad_table(int n_radial, double r0, double r1, int n_lebedev, double* out) {
    double *gx = malloc(sizeof(double) * n_radial), *gw = malloc(sizeof(double) * n_radial);
    double *lx = malloc(sizeof(double) * n_lebedev), *ly = malloc(sizeof(double) * n_lebedev);
    double *lz = malloc(sizeof(double) * n_lebedev), *lw = malloc(sizeof(double) * n_lebedev);
    int rc = 0;
    if (!gx || !gw || !lx || !ly || !lz || !lw) { rc = ORC_ERR_MEM; goto done; }
    if ((rc = orc_gauss_legendre(r0, r1, n_radial, gx, gw)) != 0) goto done;
    if ((rc = orc_lebedev(n_lebedev, lx, ly, lz, lw)) != 0) goto done;
    for (int i = 0; i < n_radial; ++i)
        for (int j = 0; j < n_lebedev; ++j) {
            double* o = out + 5 * (i * n_lebedev + j);
            o[0] = gx[i] * lx[j];
            o[1] = gx[i] * ly[j];
            o[2] = gx[i] * lz[j];
            o[3] = gx[i];
            o[4] = gw[i] * lw[j];
        }
done:
    free(gx); free(gw); free(lx); free(ly); free(lz); free(lw);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Defaults (reference constructors)                                          */
/* ------------------------------------------------------------------------- */

#define ONE_4PI_EPS0 138.935456 /* OpenMM 7.x SimTKOpenMMRealType.h */

static void common_defaults(orc_params* p) {
    memset(p, 0, sizeof(*p));
    p->n_radial = 24;
    p->r0 = 0.05;
    p->r1 = 2.0;
    p->grid_length = 0.15;
    p->sw = 0.03;
    p->rscale = 0.952;
    p->roffset = 0.03;
    p->alpha0 = -0.180;
    p->alpha1 = 1.817;
    p->eta = 0.05;
    p->slope = 1.0;
    p->shift = 0.0;
    p->vguard = 1e-30;
    p->nthreads = 1;
    p->prefactor = -ONE_4PI_EPS0 * (1.0 - 1.0 / 80.0);
}

/* CharmmReferenceGBSW.cpp:20-52 (rule 4 = 38 points, buffer 0.20, no cap) and :177-178. */
void orc_defaults_gbsw_standalone(orc_params* p) {
    common_defaults(p);
    p->model = ORC_MODEL_GBSW_STANDALONE;
    p->integral_type = ORC_INT_GBSW;
    p->n_lebedev = 38;
    p->buffer = 0.20;
    p->cap = 0;
    p->n_int = 2;
    p->orders[0] = 4;
    p->orders[1] = 7;
}

/* GBSWIntegral.cpp:24-61, GBMVIntegralTypeI.cpp:18-57, GBMVIntegralTypeII.cpp:19-62,
 * CustomGBIntegral.cpp:12-16 (cap 64, rule 3 = 26 points). */
void orc_defaults_gbmv(orc_params* p, int integral_type) {
    common_defaults(p);
    p->model = ORC_MODEL_GBMV_FORM;
    p->integral_type = integral_type;
    p->n_lebedev = 26;
    p->cap = 64;
    p->n_int = 2;
    p->orders[0] = 4;
    p->orders[1] = 7;
    if (integral_type == ORC_INT_GBSW) {
        p->buffer = 0.01 + p->sw;
    } else if (integral_type == ORC_INT_GBMV1) {
        p->buffer = 0.21;
        p->gamma0 = 0.44;
        p->beta = -100.0;
        p->lambda = 0.1;
    } else {
        p->buffer = 0.21;
        p->gamma0 = 0.44;
        p->beta = -20.0;
        p->lambda = 0.5;
        p->P1 = 0.45;
        p->P2 = 1.25;
        p->S0 = 0.7;
    }
}

/* ------------------------------------------------------------------------- */
/* Lookup grids                                                              */
/* ------------------------------------------------------------------------- */

/* d = a - b, reduced to the minimum image in a rectangular periodic box: OpenMM's
 * ReferenceForce::getDeltaRPeriodic, d -= L floor(d/L + 1/2) per axis (called with
 * (point, atom) at GBSWIntegral.cpp:168-169, CharmmReferenceGBSW.cpp:620-623 and with
 * (atom i, atom j) at CharmmReferenceGBSW.cpp:242-245, CharmmReferenceGBMV.cpp:295-296). */
static inline void delta(const orc_params* p, const double* a, const double* b, double* dx, double* dy,
                         double* dz) {
    double d[3] = {a[0] - b[0], a[1] - b[1], a[2] - b[2]};
    if (p->method == ORC_CUTOFF_PERIODIC)
        for (int k = 0; k < 3; ++k) d[k] -= p->box[k] * floor(d[k] / p->box[k] + 0.5);
    *dx = d[0];
    *dy = d[1];
    *dz = d[2];
}

/* Squared distance of atoms i and j (x_j - x_i, minimum image), summed as OpenMM's DOT3. */
static inline double pair_r2(const orc_params* p, const double* pos, int i, int j) {
    double dx, dy, dz;
    delta(p, pos + 3 * j, pos + 3 * i, &dx, &dy, &dz);
    return dx * dx + dy * dy + dz * dz;
}

/* Pair (i, j) inside the cutoff.  The standalone GBSW force skips r > c
 * (CharmmReferenceGBSW.cpp:246-247, :290-291); the CharmmGBMVForce pair terms skip r >= c
 * (CharmmReferenceGBMV.cpp:300-301).  Compared on r^2 against c^2: the same decision except
 * where sqrt rounding ties a distance to the cutoff itself. */
static inline int within_cutoff(const orc_params* p, double r2) {
    if (p->method == ORC_NO_CUTOFF) return 1;
    const double c2 = p->cutoff * p->cutoff;
    return p->model == ORC_MODEL_GBSW_STANDALONE ? !(r2 > c2) : r2 < c2;
}

typedef struct {
    double min[3], max[3], g;
    int n[3];
    long long ncell;
    int* start;  /* ncell + 1 */
    int* atoms;  /* CSR, ascending atom index per cell */
    int brute;   /* every point sees every atom */
    int N;
    int periodic;            /* cells tile the box; points are wrapped into it */
    double L[3], step[3];    /* periodic: box edges and cell edges L / n */
} grid_t;

static void grid_free(grid_t* G) {
    free(G->start);
    free(G->atoms);
    G->start = NULL;
    G->atoms = NULL;
}

/* Periodic lookup grid for a rectangular box (CustomGBIntegral.cpp:45-160): origin at the
 * centre of geometry minus half the box (:53-79), n = ceil(L/g) + 1 cells per axis of edge
 * L/n (:87-107), pad R_j + sqrt3/2 max(L/n) + buffer (:114-119).  GBMV form: the reference's
 * membership, an atom-anchored sample lattice walked with floating accumulation (:120-157,
 * below).  ORC_POLICY_CORNERS instead applies the non-periodic corner rule at the image of
 * each cell nearest the atom (atom j, wrapped into the box, joins the cells of the unwrapped
 * index range around it, each index reduced modulo n), kept only to measure how much the
 * choice moves the result.  Atoms in ascending index, cap in index order.  The standalone
 * GBSW model uses its one pad (CharmmReferenceGBSW.cpp:454-455) with the cell edge max(L/n)
 * and the corner rule; its own periodic lookup keeps the non-periodic grid and shifts only
 * points outside it (:527-536), which misses the images of atoms across the box faces. */
static int grid_build_periodic(grid_t* G, const orc_params* p, int N, const double* pos, const double* R) {
    const int standalone = (p->model == ORC_MODEL_GBSW_STANDALONE);
    const double g = p->grid_length;
    double cog[3] = {0.0, 0.0, 0.0}, maxR = 0.0, maxstep = 0.0;
    for (int a = 0; a < N; ++a) {
        for (int d = 0; d < 3; ++d) cog[d] += pos[3 * a + d];
        if (R[a] > maxR) maxR = R[a];
    }
    long long ncell = 1;
    for (int d = 0; d < 3; ++d) {
        cog[d] /= N;
        const double L = p->box[d];
        G->L[d] = L;
        G->n[d] = (int)ceil(L / g) + 1;
        G->step[d] = L / G->n[d];
        G->min[d] = cog[d] - L / 2.0;
        G->max[d] = G->min[d] + L;
        if (G->step[d] > maxstep) maxstep = G->step[d];
        ncell *= G->n[d];
    }
    G->periodic = 1;
    G->g = maxstep;
    G->ncell = ncell;
    const double halfdiag = sqrt(3.0) / 2.0 * maxstep;
    const double gpad = p->buffer + maxR + p->sw + halfdiag + 1e-6;
    /* every cell's list must hold one image of an atom at most, and that image must be the
     * nearest to every point of the cell */
    const double padmax = standalone ? gpad : maxR + halfdiag + p->buffer;
    for (int d = 0; d < 3; ++d)
        if (!(G->L[d] > 2.0 * (padmax + sqrt(3.0) * maxstep))) return ORC_ERR_ARG;
    const int cap = p->cap > 0 ? p->cap : INT32_MAX;
    int* count = calloc((size_t)ncell, sizeof(int));
    G->start = calloc((size_t)ncell + 1, sizeof(int));
    if (!count || !G->start) { free(count); return ORC_ERR_MEM; }
    const int nx = G->n[0], ny = G->n[1], nz = G->n[2];
    if (p->policy != ORC_POLICY_CORNERS && !standalone) {
        /* The reference's own membership (CustomGBIntegral.cpp:114-157): for each atom in
         * index order, sample points walked from x - pad - g' to x + pad + g' (g' = max L/n)
         * in steps of L/n per axis with floating accumulation; each sample is wrapped into the
         * box for its cell (:135-140) and kept when the SAMPLE lies within pad of the atom
         * (:143-145), appended while the cell holds fewer than cap atoms (:149-154).  Repeats of
         * an atom in a cell are kept, as the reference keeps them.  Cell indices are clamped to
         * [0, n-1] (a wrapped coordinate that rounds to the upper box face would index past the
         * grid in the reference). */
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) {
                long long s = 0;
                for (long long c = 0; c < ncell; ++c) { G->start[c] = (int)s; s += count[c]; }
                G->start[ncell] = (int)s;
                G->atoms = malloc(sizeof(int) * (size_t)(s > 0 ? s : 1));
                if (!G->atoms) { free(count); return ORC_ERR_MEM; }
                memset(count, 0, sizeof(int) * (size_t)ncell);
            }
            for (int a = 0; a < N; ++a) {
                const double* co = pos + 3 * a;
                const double pad = R[a] + sqrt(3.0) / 2.0 * maxstep + p->buffer;
                double bc[3], ec[3];
                for (int d = 0; d < 3; ++d) {
                    bc[d] = co[d] - pad - maxstep;
                    ec[d] = co[d] + pad + maxstep;
                }
                for (double x = bc[0]; x < ec[0]; x += G->step[0])
                    for (double y = bc[1]; y < ec[1]; y += G->step[1])
                        for (double z = bc[2]; z < ec[2]; z += G->step[2]) {
                            const double smp[3] = {x, y, z};
                            int idx[3];
                            for (int d = 0; d < 3; ++d) {
                                const double w = smp[d] - G->L[d] * floor((smp[d] - G->min[d]) / G->L[d]);
                                idx[d] = (int)floor((w - G->min[d]) / G->step[d]);
                                if (idx[d] < 0) idx[d] = 0;
                                if (idx[d] > G->n[d] - 1) idx[d] = G->n[d] - 1;
                            }
                            const double dx = x - co[0], dy = y - co[1], dz = z - co[2];
                            if (!(dx * dx + dy * dy + dz * dz < pad * pad)) continue;
                            const long long c = ((long long)idx[0] * ny + idx[1]) * nz + idx[2];
                            if (count[c] < cap) {
                                if (pass == 1) G->atoms[G->start[c] + count[c]] = a;
                                count[c]++;
                            }
                        }
            }
        }
        free(count);
        return 0;
    }
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            long long s = 0;
            for (long long c = 0; c < ncell; ++c) { G->start[c] = (int)s; s += count[c]; }
            G->start[ncell] = (int)s;
            G->atoms = malloc(sizeof(int) * (size_t)(s > 0 ? s : 1));
            if (!G->atoms) { free(count); return ORC_ERR_MEM; }
            memset(count, 0, sizeof(int) * (size_t)ncell);
        }
        for (int a = 0; a < N; ++a) {
            double xw[3];
            for (int d = 0; d < 3; ++d)
                xw[d] = pos[3 * a + d] - G->L[d] * floor((pos[3 * a + d] - G->min[d]) / G->L[d]);
            const double pad = standalone ? gpad : R[a] + halfdiag + p->buffer;
            int b[3], e[3];
            for (int d = 0; d < 3; ++d) {
                b[d] = (int)floor((xw[d] - pad - G->min[d]) / G->step[d]);
                e[d] = (int)ceil((xw[d] + pad - G->min[d]) / G->step[d]);
            }
            for (int i = b[0]; i <= e[0]; ++i)
                for (int j = b[1]; j <= e[1]; ++j)
                    for (int k = b[2]; k <= e[2]; ++k) {
                        double dx = (G->min[0] + i * G->step[0]) - xw[0];
                        double dy = (G->min[1] + j * G->step[1]) - xw[1];
                        double dz = (G->min[2] + k * G->step[2]) - xw[2];
                        int member;
                        if (standalone)
                            member = sqrt(dx * dx + dy * dy + dz * dz) < pad;
                        else
                            member = dx * dx + dy * dy + dz * dz < pad * pad;
                        if (!member) continue;
                        const int iw = ((i % nx) + nx) % nx, jw = ((j % ny) + ny) % ny, kw = ((k % nz) + nz) % nz;
                        long long c = ((long long)iw * ny + jw) * nz + kw;
                        if (count[c] < cap) {
                            if (pass == 1) G->atoms[G->start[c] + count[c]] = a;
                            count[c]++;
                        }
                    }
        }
    }
    free(count);
    return 0;
}

/* Atom j joins cell (i,j,k) when the cell's lower corner lies within pad of x_j;
 * atoms are visited in ascending index and a cell keeps the first `cap`.
 * GBMV form: CustomGBIntegral.cpp:162-228 (grid pad 0.03 + sqrt3/2 g + buffer,
 *            per-atom pad R_j + sqrt3/2 g + buffer, squared comparison).
 * Standalone: CharmmReferenceGBSW.cpp:440-514 (one pad buffer + maxR + sw + sqrt3/2 g
 *            + 1e-6 for grid and atoms, sqrt comparison, no cap). */
static int grid_build(grid_t* G, const orc_params* p, int N, const double* pos, const double* R) {
    memset(G, 0, sizeof(*G));
    G->N = N;
    if (p->policy == ORC_POLICY_BRUTE) {
        G->brute = 1;
        return 0;
    }
    if (p->method == ORC_CUTOFF_PERIODIC) return grid_build_periodic(G, p, N, pos, R);
    const int standalone = (p->model == ORC_MODEL_GBSW_STANDALONE);
    const double g = p->grid_length;
    double mn[3], mx[3];
    for (int d = 0; d < 3; ++d) mn[d] = mx[d] = pos[d];
    double maxR = 0.0;
    for (int a = 0; a < N; ++a) {
        for (int d = 0; d < 3; ++d) {
            if (pos[3 * a + d] < mn[d]) mn[d] = pos[3 * a + d];
            if (pos[3 * a + d] > mx[d]) mx[d] = pos[3 * a + d];
        }
        if (R[a] > maxR) maxR = R[a];
    }
    double gpad;
    if (standalone)
        gpad = p->buffer + maxR + p->sw + sqrt(3.0) / 2.0 * g + 1e-6;
    else
        gpad = 0.03 + sqrt(3.0) / 2.0 * g + p->buffer;
    long long ncell = 1;
    for (int d = 0; d < 3; ++d) {
        mn[d] -= gpad;
        mx[d] += gpad;
        double len = mx[d] - mn[d];
        if (len > 1000.0) return ORC_ERR_ARG; /* CustomGBIntegral.cpp:182-183 */
        G->n[d] = (int)ceil(len / g) + 1;
        G->min[d] = mn[d];
        G->max[d] = mn[d] + (G->n[d] - 1) * g;
        ncell *= G->n[d];
    }
    G->g = g;
    G->ncell = ncell;
    const int cap = p->cap > 0 ? p->cap : INT32_MAX;
    int* count = calloc((size_t)ncell, sizeof(int));
    G->start = calloc((size_t)ncell + 1, sizeof(int));
    if (!count || !G->start) { free(count); return ORC_ERR_MEM; }
    const int ny = G->n[1], nz = G->n[2];
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            long long s = 0;
            for (long long c = 0; c < ncell; ++c) { G->start[c] = (int)s; s += count[c]; }
            G->start[ncell] = (int)s;
            G->atoms = malloc(sizeof(int) * (size_t)(s > 0 ? s : 1));
            if (!G->atoms) { free(count); return ORC_ERR_MEM; }
            memset(count, 0, sizeof(int) * (size_t)ncell);
        }
        for (int a = 0; a < N; ++a) {
            const double* x = pos + 3 * a;
            double pad = standalone ? gpad : R[a] + sqrt(3.0) / 2.0 * g + p->buffer;
            int b[3], e[3];
            for (int d = 0; d < 3; ++d) {
                b[d] = (int)floor((x[d] - pad - G->min[d]) / g);
                e[d] = (int)ceil((x[d] + pad - G->min[d]) / g);
                if (b[d] < 0) b[d] = 0;
                if (e[d] > G->n[d] - 1) e[d] = G->n[d] - 1;
            }
            for (int i = b[0]; i <= e[0]; ++i)
                for (int j = b[1]; j <= e[1]; ++j)
                    for (int k = b[2]; k <= e[2]; ++k) {
                        double dx = (G->min[0] + i * g) - x[0];
                        double dy = (G->min[1] + j * g) - x[1];
                        double dz = (G->min[2] + k * g) - x[2];
                        int member;
                        if (standalone)
                            member = sqrt(dx * dx + dy * dy + dz * dz) < pad;
                        else
                            member = dx * dx + dy * dy + dz * dz < pad * pad;
                        if (!member) continue;
                        long long c = ((long long)i * ny + j) * nz + k;
                        if (count[c] < cap) {
                            if (pass == 1) G->atoms[G->start[c] + count[c]] = a;
                            count[c]++;
                        }
                    }
        }
    }
    free(count);
    return 0;
}

/* Point -> cell: outside -> -1; else floor((p - min)/g) per axis
 * (CustomGBIntegral.cpp:310-324, CharmmReferenceGBSW.cpp:538-549). */
static long long grid_cell(const grid_t* G, const double* pt) {
    int idx[3];
    if (G->periodic) {
        /* the point wrapped into the box (CustomGBIntegral.cpp:294-301) */
        for (int d = 0; d < 3; ++d) {
            const double pw = pt[d] - floor((pt[d] - G->min[d]) / G->L[d]) * G->L[d];
            idx[d] = (int)floor((pw - G->min[d]) / G->step[d]);
            if (idx[d] < 0) idx[d] = 0;
            if (idx[d] > G->n[d] - 1) idx[d] = G->n[d] - 1;
        }
        return ((long long)idx[0] * G->n[1] + idx[1]) * G->n[2] + idx[2];
    }
    for (int d = 0; d < 3; ++d) {
        if (pt[d] < G->min[d] || pt[d] > G->max[d]) return -1;
        idx[d] = (int)floor((pt[d] - G->min[d]) / G->g);
    }
    return ((long long)idx[0] * G->n[1] + idx[1]) * G->n[2] + idx[2];
}

/* Neighbour list of a point. */
static int point_list(const grid_t* G, const double* pt, const int** list, int* brute_scratch) {
    if (G->brute) {
        *list = brute_scratch;
        return G->N;
    }
    long long c = grid_cell(G, pt);
    if (c < 0) { *list = NULL; return 0; }
    *list = G->atoms + G->start[c];
    return G->start[c + 1] - G->start[c];
}

/* ------------------------------------------------------------------------- */
/* Per-point volume functions and their gradients                             */
/* ------------------------------------------------------------------------- */

/* GBSW smoothing H(d) = 1/2 + 3/(4w) x - x^3/(4w^3), x = d - R', and dH/dd. */
static inline double gbsw_h(double x, double w) {
    return 0.5 + 3.0 / (4.0 * w) * x - 1.0 / (4.0 * w * w * w) * x * x * x;
}
static inline double gbsw_dh(double x, double w) {
    return 3.0 / (4.0 * w) - 3.0 / (4.0 * w * w * w) * x * x;
}

/* V = prod_j H_j over the list in list order; 0 as soon as a point is buried
 * (GBSWIntegral.cpp:187-224, CharmmReferenceGBSW.cpp:616-639). */
static double gbsw_volume(const orc_params* p, const double* pos, const double* R, const double* pt,
                          const int* list, int n) {
    double V = 1.0;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double Rj = (R[j] + p->roffset) * p->rscale;
        if (r <= Rj - p->sw) return 0.0;
        if (r >= Rj + p->sw) continue;
        double h = gbsw_h(r - Rj, p->sw);
        if (h <= 0.0) return 0.0;
        V *= h;
    }
    return V;
}

/* dI/dx for I += pref (1 - V): g_j = -pref V H'_j/H_j (x_j - p)/d for shell atoms j != owner,
 * g_owner = -sum_j g_j (GBSWIntegral.cpp:271-314; CharmmReferenceGBSW.cpp:641-674).
 * Adds scale * g to out (N x 3).  trunc: the standalone force with a cutoff applies
 * dB_i/dx_j only to atoms j within the cutoff of the owner i, while the owner keeps the
 * whole sum (CharmmReferenceGBSW.cpp:285-296 skips r_ij > c; dB_i/dx_i, r = 0, collects
 * every j, :670-671). */
static int gbsw_gradient(const orc_params* p, const double* pos, const double* R, const double* pt,
                         int owner, const int* list, int n, double V, double pref, double* out, int trunc) {
    if (V == 0.0) return 0;
    int m = 0;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        if (j == owner) continue;
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double Rj = (R[j] + p->roffset) * p->rscale;
        if (!(r > Rj - p->sw && r < Rj + p->sw)) continue;
        double x = r - Rj;
        double h = gbsw_h(x, p->sw), dh = gbsw_dh(x, p->sw);
        double f = pref * (dh * V / h) / r;
        out[3 * owner] += dx * f;
        out[3 * owner + 1] += dy * f;
        out[3 * owner + 2] += dz * f;
        ++m;
        if (trunc && !within_cutoff(p, pair_r2(p, pos, owner, j))) continue;
        out[3 * j] -= dx * f;
        out[3 * j + 1] -= dy * f;
        out[3 * j + 2] -= dz * f;
    }
    return m;
}

/* GBMV-I: S = sum_j exp(gamma_j d^4), gamma_j = gamma0 ln(lambda)/R_j^4; returns
 * pre = exp(beta (S - lambda)), V = 1/(1+pre) (GBMVIntegralTypeI.cpp:155-173). */
static double gbmv1_pre(const orc_params* p, const double* pos, const double* R, const double* pt,
                        const int* list, int n) {
    double S = 0.0;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double r4 = r * r * r * r;
        double R4 = R[j] * R[j] * R[j] * R[j];
        double gam = p->gamma0 * log(p->lambda) / R4;
        S += exp(gam * r4);
    }
    return exp(p->beta * (S - p->lambda));
}

/* GBMVIntegralTypeI.cpp:195-224: dI/dx for I += pref V. */
static int gbmv1_gradient(const orc_params* p, const double* pos, const double* R, const double* pt,
                          int owner, const int* list, int n, double pre, double pref, double* out) {
    double factor = -1.0 / ((1.0 + pre) * (1.0 + pre)) * (p->beta * pre) * pref;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        if (j == owner) continue; /* its two terms cancel in the reference */
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double r3 = r * r * r, r4 = r3 * r;
        double R4 = R[j] * R[j] * R[j] * R[j];
        double gam = p->gamma0 * log(p->lambda) / R4;
        double f2 = factor * gam * exp(gam * r4) * (4.0 * r3) / r;
        out[3 * owner] -= dx * f2;
        out[3 * owner + 1] -= dy * f2;
        out[3 * owner + 2] -= dz * f2;
        out[3 * j] += dx * f2;
        out[3 * j + 1] += dy * f2;
        out[3 * j + 2] += dz * f2;
    }
    return n;
}

typedef struct {
    double sum1, sum2, vsq, v[3];
    double pre;   /* exp(beta (S - lambda)); 0 when guarded */
    int guarded;
} gbmv2_sums;

/* GBMV-II (GBMVIntegralTypeII.cpp:163-191): C_j = 10 P1 R_j + P2,
 * F_j = (C_j / (C_j + 100 d^2 - 100 R_j^2))^2, sum1 = sum F, sum2 = sum d^2 F^2,
 * v = sum F (x_j - p), S = S0 sum1 sum2 / |v|^2, pre = exp(beta (S - lambda)). */
static void gbmv2_point(const orc_params* p, const double* pos, const double* R, const double* pt,
                        const int* list, int n, gbmv2_sums* s) {
    double sum1 = 0.0, sum2 = 0.0, v0 = 0.0, v1 = 0.0, v2 = 0.0;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double d2 = dx * dx + dy * dy + dz * dz;
        double C = p->P1 * R[j] * 10.0 + p->P2;
        double F = C / (C + d2 * 100.0 - R[j] * R[j] * 100.0);
        F = F * F;
        sum1 += F;
        sum2 += d2 * (F * F);
        v0 += dx * F;
        v1 += dy * F;
        v2 += dz * F;
    }
    s->sum1 = sum1;
    s->sum2 = sum2;
    s->v[0] = v0;
    s->v[1] = v1;
    s->v[2] = v2;
    s->vsq = v0 * v0 + v1 * v1 + v2 * v2;
    s->guarded = !(s->vsq >= p->vguard);
    if (s->guarded) {
        s->pre = 0.0;
    } else {
        double S = p->S0 * sum1 * sum2 / s->vsq;
        s->pre = exp(p->beta * (S - p->lambda));
    }
}

/* GBMVIntegralTypeII.cpp:221-282: dI/dx for I += pref V, three terms per list atom. */
static int gbmv2_gradient(const orc_params* p, const double* pos, const double* R, const double* pt,
                          int owner, const int* list, int n, const gbmv2_sums* s, double pref,
                          double* out) {
    if (s->guarded) return 0;
    const double pre = s->pre;
    const double factor = -1.0 / ((1.0 + pre) * (1.0 + pre)) * (p->beta * pre) * pref * p->S0;
    const double vsq = s->vsq;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        if (j == owner) continue; /* its terms cancel in the reference */
        double dx, dy, dz;
        delta(p, pos + 3 * j, pt, &dx, &dy, &dz);
        double d2 = dx * dx + dy * dy + dz * dz;
        double C = p->P1 * R[j] * 10.0 + p->P2;
        double t = C + d2 * 100.0 - R[j] * R[j] * 100.0;
        double F = C / t;
        F = F * F;
        double D = -400.0 * C * C / (t * t * t);
        double c1 = D * s->sum2 / vsq * factor;
        double c2 = 2.0 * F * (F + D * d2) / vsq * s->sum1 * factor;
        double c3 = -2.0 / (vsq * vsq) * s->sum1 * s->sum2 * factor;
        double vdr = s->v[0] * dx + s->v[1] * dy + s->v[2] * dz;
        double gx = (c1 + c2) * dx + c3 * (s->v[0] * F + vdr * D * dx);
        double gy = (c1 + c2) * dy + c3 * (s->v[1] * F + vdr * D * dy);
        double gz = (c1 + c2) * dz + c3 * (s->v[2] * F + vdr * D * dz);
        out[3 * owner] -= gx;
        out[3 * owner + 1] -= gy;
        out[3 * owner + 2] -= gz;
        out[3 * j] += gx;
        out[3 * j + 1] += gy;
        out[3 * j + 2] += gz;
    }
    return n - 1;
}

/* ------------------------------------------------------------------------- */
/* Born radius map and STILL energy                                           */
/* ------------------------------------------------------------------------- */

/* B = slope/D + shift, D = a0 (1/eta - I4) + a1 (1/(4 eta^4) - I7)^(1/4)
 * (CharmmReferenceGBSW.cpp:591-593 with slope 1, shift 0); dB/dI4, dB/dI7
 * (CharmmReferenceGBSW.cpp:607-609). */
static void born_map(const orc_params* p, double I4, double I7, double* B, double* dB4, double* dB7) {
    double eta4 = p->eta * p->eta * p->eta * p->eta;
    double T = 1.0 / (4.0 * eta4) - I7;
    double D = p->alpha0 * (1.0 / p->eta - I4) + p->alpha1 * pow(T, 0.25);
    *B = p->slope / D + p->shift;
    *dB4 = p->slope * p->alpha0 / (D * D);
    *dB7 = p->slope * 0.25 * p->alpha1 * pow(T, -0.75) / (D * D);
}

/* STILL pair + self energy (CharmmReferenceGBSW.cpp:237-283):
 * E = sum_i P q_i^2/(2 B_i) + sum_{i<j} P q_i q_j / f_ij,
 * f_ij = sqrt(r^2 + B_i B_j exp(-r^2/(4 B_i B_j))).  Adds pair forces to F,
 * overwrites dEdB.  With a cutoff only pairs inside it count (minimum image when
 * periodic); the self term always does (CharmmReferenceGBSW.cpp:188-199: every atom
 * pairs with itself; CharmmReferenceGBMV.cpp: it is a SingleParticle term). */
static int cmp_ll(const void* a, const void* b) {
    const long long x = *(const long long*)a, y = *(const long long*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static double still_energy(const orc_params* p, int N, const double* pos, const double* q, const double* B,
                           double P, double* F, double* dEdB) {
    double E = 0.0;
    /* excluded pairs as sorted keys i N + j (i < j); CharmmReferenceGBMV.cpp:272-284 skips
     * them in a ParticlePair term */
    long long* ex = NULL;
    int nex = 0;
    if (p->model == ORC_MODEL_GBMV_FORM && p->n_excl > 0 && p->excl) {
        ex = malloc(sizeof(long long) * (size_t)p->n_excl);
        for (int k = 0; k < p->n_excl && ex; ++k) {
            int a = p->excl[2 * k], b = p->excl[2 * k + 1];
            if (a > b) { int t = a; a = b; b = t; }
            ex[nex++] = (long long)a * N + b;
        }
        if (ex) qsort(ex, (size_t)nex, sizeof(long long), cmp_ll);
    }
    for (int i = 0; i < N; ++i) dEdB[i] = 0.0;
    for (int i = 0; i < N; ++i) {
        E += 0.5 * P * q[i] * q[i] / B[i];
        dEdB[i] += -0.5 * P * q[i] * q[i] / (B[i] * B[i]);
        for (int j = i + 1; j < N; ++j) {
            double dx, dy, dz;
            delta(p, pos + 3 * j, pos + 3 * i, &dx, &dy, &dz);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (!within_cutoff(p, r2)) continue;
            if (ex) {
                const long long key = (long long)i * N + j;
                if (bsearch(&key, ex, (size_t)nex, sizeof(long long), cmp_ll)) continue;
            }
            double BB = B[i] * B[j];
            double D = r2 / (4.0 * BB);
            double ex = exp(-D);
            double f = sqrt(r2 + BB * ex);
            double f3 = f * f * f;
            double qq = q[i] * q[j];
            E += P * qq / f;
            double dGdr = -0.25 * P * qq * (4.0 - ex) / f3;
            F[3 * i] += dx * dGdr;
            F[3 * i + 1] += dy * dGdr;
            F[3 * i + 2] += dz * dGdr;
            F[3 * j] -= dx * dGdr;
            F[3 * j + 1] -= dy * dGdr;
            F[3 * j + 2] -= dz * dGdr;
            double common = -0.5 * P * qq * ex / f3;
            dEdB[i] += common * (B[j] + r2 / (4.0 * B[i]));
            dEdB[j] += common * (B[i] + r2 / (4.0 * B[j]));
        }
    }
    free(ex);
    return E;
}

/* ------------------------------------------------------------------------- */
/* Driver                                                                     */
/* ------------------------------------------------------------------------- */

typedef struct {
    const orc_params* p;
    int N, Q, nint;
    const double* pos;
    const double* R;
    const double* quad;  /* Q x 5 */
    double* pw;          /* nint x Q : w_q / r_q^(order-2) */
    grid_t G;
    int* brute_list;
    int trunc;  /* standalone GBSW with a cutoff: truncated chain rule */
} ctx_t;

/* V of one point and its list; kind-specific sums in *s2 / *pre. */
static double point_volume(const ctx_t* c, const double* pt, const int* list, int n, gbmv2_sums* s2,
                           double* pre) {
    const orc_params* p = c->p;
    switch (p->integral_type) {
    case ORC_INT_GBSW:
        return n == 0 ? 1.0 : gbsw_volume(p, c->pos, c->R, pt, list, n);
    case ORC_INT_GBMV1: {
        *pre = gbmv1_pre(p, c->pos, c->R, pt, list, n);
        return 1.0 / (1.0 + *pre);
    }
    default: {
        if (n == 0) { s2->guarded = 1; s2->pre = 0.0; return 0.0; }
        gbmv2_point(p, c->pos, c->R, pt, list, n, s2);
        /* guarded: the S -> +inf limit of 1/(1 + exp(beta (S - lambda))) */
        if (s2->guarded) return p->beta < 0.0 ? 1.0 : 0.0;
        return 1.0 / (1.0 + s2->pre);
    }
    }
}

/* Contribution of one point to integral k: GBSW accumulates 1-V, GBMV accumulates V. */
static inline double point_integrand(int type, double V) { return type == ORC_INT_GBSW ? 1.0 - V : V; }

static int point_gradient(const ctx_t* c, const double* pt, int owner, const int* list, int n,
                          double V, const gbmv2_sums* s2, double pre, double pref, double* out) {
    const orc_params* p = c->p;
    switch (p->integral_type) {
    case ORC_INT_GBSW:
        return n == 0 ? 0 : gbsw_gradient(p, c->pos, c->R, pt, owner, list, n, V, pref, out, c->trunc);
    case ORC_INT_GBMV1:
        return gbmv1_gradient(p, c->pos, c->R, pt, owner, list, n, pre, pref, out);
    default:
        return n == 0 ? 0 : gbmv2_gradient(p, c->pos, c->R, pt, owner, list, n, s2, pref, out);
    }
}

static int gbsw_exact_count(const ctx_t* c, const double* pt, const int* list, int n) {
    int m = 0;
    for (int k = 0; k < n; ++k) {
        int j = list[k];
        double dx, dy, dz;
        delta(c->p, c->pos + 3 * j, pt, &dx, &dy, &dz);
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        double Rj = (c->R[j] + c->p->roffset) * c->p->rscale;
        if (r < Rj + c->p->sw) ++m;
    }
    return m;
}

int orc_compute(const orc_params* p, int N, const double* pos, const double* charge,
                const double* radius, double* energy, double* forces, double* born,
                double* integrals_out, double* dEdI_out, orc_stats* stats) {
    if (!p || N <= 0 || !pos || !charge || !radius || !forces || !energy) return ORC_ERR_ARG;
    const int standalone = (p->model == ORC_MODEL_GBSW_STANDALONE);
    const int type = standalone ? ORC_INT_GBSW : p->integral_type;
    orc_params pp = *p;
    pp.integral_type = type;
    if (standalone) { pp.n_int = 2; pp.orders[0] = 4; pp.orders[1] = 7; }
    if (pp.n_int != 2) return ORC_ERR_ARG;
    int i4 = -1, i7 = -1;
    for (int k = 0; k < 2; ++k) {
        if (pp.orders[k] == 4) i4 = k;
        if (pp.orders[k] == 7) i7 = k;
    }
    if (i4 < 0 || i7 < 0) return ORC_ERR_ARG; /* canonical Born map needs orders 4 and 7 */

    if (pp.method < ORC_NO_CUTOFF || pp.method > ORC_CUTOFF_PERIODIC) return ORC_ERR_ARG;
    if (pp.n_excl < 0 || (pp.n_excl > 0 && !pp.excl)) return ORC_ERR_ARG;
    for (int k = 0; k < pp.n_excl; ++k) {
        const int a = pp.excl[2 * k], b = pp.excl[2 * k + 1];
        if (a < 0 || b < 0 || a >= N || b >= N || a == b) return ORC_ERR_ARG;
    }
    if (pp.method != ORC_NO_CUTOFF && !(pp.cutoff > 0.0)) return ORC_ERR_ARG;
    if (pp.method == ORC_CUTOFF_PERIODIC)
        for (int d = 0; d < 3; ++d) /* CharmmGBMVForceImpl.cpp:99-104, CharmmGBSWForceImpl.cpp:48-54 */
            if (!(pp.box[d] >= 2.0 * pp.cutoff)) return ORC_ERR_ARG;

    ctx_t c;
    memset(&c, 0, sizeof(c));
    c.trunc = standalone && pp.method != ORC_NO_CUTOFF;
    c.p = &pp;
    c.N = N;
    c.pos = pos;
    c.R = radius;
    c.nint = 2;
    c.Q = pp.n_radial * pp.n_lebedev;
    const int Q = c.Q;
    double* quad = malloc(sizeof(double) * 5 * Q);
    c.pw = malloc(sizeof(double) * 2 * Q);
    double* I = calloc((size_t)2 * N, sizeof(double));
    double* B = malloc(sizeof(double) * N);
    double* dB = malloc(sizeof(double) * 2 * N);
    double* dEdB = malloc(sizeof(double) * N);
    double* dEdI = malloc(sizeof(double) * 2 * N);
    c.brute_list = malloc(sizeof(int) * N);
    int rc = 0;
    if (!quad || !c.pw || !I || !B || !dB || !dEdB || !dEdI || !c.brute_list) { rc = ORC_ERR_MEM; goto out; }
    for (int a = 0; a < N; ++a) c.brute_list[a] = a;
    if ((rc = orc_quad_table(pp.n_radial, pp.r0, pp.r1, pp.n_lebedev, quad)) != 0) goto out;
    c.quad = quad;
    for (int k = 0; k < 2; ++k)
        for (int q = 0; q < Q; ++q) c.pw[k * Q + q] = quad[5 * q + 4] / pow(quad[5 * q + 3], pp.orders[k] - 2);
    if ((rc = grid_build(&c.G, &pp, N, pos, radius)) != 0) goto out;

    memset(forces, 0, sizeof(double) * 3 * N);
    double n_points = 0, n_vol = 0, n_exact = 0, n_frc = 0;

    /* ---- integrals (+ dense gradients) ---- */
    double* dense = NULL; /* standalone: N x N x 3 (dB_i/dx_j); GBMV: 2 x N x N x 3 (dI_k(i)/dx_j) */
    if (pp.dense) {
        size_t len = (size_t)(standalone ? 1 : 2) * N * N * 3;
        dense = calloc(len, sizeof(double));
        if (!dense) { rc = ORC_ERR_MEM; goto out; }
    }
    if (pp.dense && !standalone) {
        /* reference order: q outer, atom inner (GBSWIntegral.cpp:128-131) */
        for (int q = 0; q < Q; ++q) {
            for (int i = 0; i < N; ++i) {
                double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                pos[3 * i + 2] + quad[5 * q + 2]};
                const int* list;
                int n = point_list(&c.G, pt, &list, c.brute_list);
                if (type == ORC_INT_GBSW && n == 0) continue;
                gbmv2_sums s2;
                double pre = 0;
                double V = point_volume(&c, pt, list, n, &s2, &pre);
                for (int k = 0; k < 2; ++k) I[k * N + i] += c.pw[k * Q + q] * point_integrand(type, V);
                for (int k = 0; k < 2; ++k)
                    point_gradient(&c, pt, i, list, n, V, &s2, pre, c.pw[k * Q + q],
                                   dense + ((size_t)k * N + i) * N * 3);
            }
        }
    } else {
        /* per atom, q in order (deterministic); standalone dense keeps its own second loop */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(pp.nthreads > 1 ? pp.nthreads : 1) if (!pp.dense)
#endif
        for (int i = 0; i < N; ++i) {
            double acc[2] = {0, 0};
            for (int q = 0; q < Q; ++q) {
                double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                pos[3 * i + 2] + quad[5 * q + 2]};
                const int* list;
                int n = point_list(&c.G, pt, &list, c.brute_list);
                if (type == ORC_INT_GBSW && n == 0) continue;
                gbmv2_sums s2;
                double pre = 0;
                double V = point_volume(&c, pt, list, n, &s2, &pre);
                for (int k = 0; k < 2; ++k) acc[k] += c.pw[k * Q + q] * point_integrand(type, V);
            }
            I[i] = acc[0];
            I[N + i] = acc[1];
        }
    }

    /* ---- Born radii ---- */
    for (int i = 0; i < N; ++i) {
        double b4, b7;
        born_map(&pp, I[i4 * N + i], I[i7 * N + i], &B[i], &b4, &b7);
        dB[i4 * N + i] = b4;
        dB[i7 * N + i] = b7;
    }
    if (pp.dense && standalone) {
        /* compute_dbornR_dr_vec_Fast with the combined prefactor (CharmmReferenceGBSW.cpp:597-611) */
        for (int i = 0; i < N; ++i) {
            for (int q = 0; q < Q; ++q) {
                double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                pos[3 * i + 2] + quad[5 * q + 2]};
                const int* list;
                int n = point_list(&c.G, pt, &list, c.brute_list);
                if (n == 0) continue;
                double V = gbsw_volume(&pp, pos, radius, pt, list, n);
                double pref = dB[i4 * N + i] * c.pw[i4 * Q + q] + dB[i7 * N + i] * c.pw[i7 * Q + q];
                /* gbsw_gradient returns dI/dx for I += pref(1-V); dB/dx has the same form */
                gbsw_gradient(&pp, pos, radius, pt, i, list, n, V, pref, dense + (size_t)i * N * 3, 0);
            }
        }
    }

    /* ---- pair + self energy ---- */
    double E = still_energy(&pp, N, pos, charge, B, pp.prefactor, forces, dEdB);
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < N; ++i) dEdI[k * N + i] = dEdB[i] * dB[k * N + i];

    /* ---- chain rule ---- */
    if (pp.dense) {
        if (standalone) {
            /* CharmmReferenceGBSW.cpp:285-296: with a cutoff, j beyond it of i is skipped */
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j) {
                    if (c.trunc && j != i && !within_cutoff(&pp, pair_r2(&pp, pos, i, j))) continue;
                    for (int d = 0; d < 3; ++d) forces[3 * j + d] -= dEdB[i] * dense[((size_t)i * N + j) * 3 + d];
                }
        } else {
            for (int i = 0; i < N; ++i)
                for (int j = 0; j < N; ++j)
                    for (int k = 0; k < 2; ++k)
                        for (int d = 0; d < 3; ++d)
                            forces[3 * j + d] -= dEdI[k * N + i] * dense[(((size_t)k * N + i) * N + j) * 3 + d];
        }
    } else {
        int nt = 1;
#ifdef _OPENMP
        nt = pp.nthreads > 1 ? pp.nthreads : 1;
#endif
        double* part = calloc((size_t)nt * N * 3, sizeof(double));
        if (!part) { rc = ORC_ERR_MEM; free(dense); goto out; }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            double* g = part + (size_t)tid * N * 3;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int i = 0; i < N; ++i) {
                for (int q = 0; q < Q; ++q) {
                    double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                    pos[3 * i + 2] + quad[5 * q + 2]};
                    const int* list;
                    int n = point_list(&c.G, pt, &list, c.brute_list);
                    if (type == ORC_INT_GBSW && n == 0) continue;
                    gbmv2_sums s2;
                    double pre = 0;
                    double V = point_volume(&c, pt, list, n, &s2, &pre);
                    double pref = dEdI[i] * c.pw[q] + dEdI[N + i] * c.pw[Q + q];
                    point_gradient(&c, pt, i, list, n, V, &s2, pre, pref, g);
                }
            }
        }
        for (int t = 0; t < nt; ++t)
            for (int a = 0; a < 3 * N; ++a) forces[a] -= part[(size_t)t * N * 3 + a];
        free(part);
    }
    free(dense);

    /* ---- statistics (serial; only when asked) ---- */
    if (stats) {
        double* scratch = calloc((size_t)3 * N, sizeof(double));
        for (int i = 0; i < N && scratch; ++i) {
            for (int q = 0; q < Q; ++q) {
                double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                pos[3 * i + 2] + quad[5 * q + 2]};
                const int* list;
                int n = point_list(&c.G, pt, &list, c.brute_list);
                if (n == 0) continue;
                n_points += 1;
                n_vol += n;
                gbmv2_sums s2;
                double pre = 0;
                double V = point_volume(&c, pt, list, n, &s2, &pre);
                if (type == ORC_INT_GBSW) {
                    n_exact += gbsw_exact_count(&c, pt, list, n);
                    if (V > 0.0 && V < 1.0) n_frc += point_gradient(&c, pt, i, list, n, V, &s2, pre, 0.0, scratch);
                } else {
                    n_frc += n;
                }
            }
        }
        free(scratch);
        stats->energy = E;
        stats->n_points = n_points;
        stats->n_vol_pairs = n_vol;
        stats->n_vol_exact = type == ORC_INT_GBSW ? n_exact : n_vol;
        stats->n_frc_pairs = n_frc;
        stats->n_cells = (double)c.G.ncell;
        stats->n_entries = c.G.brute ? (double)N : (double)(c.G.start ? c.G.start[c.G.ncell] : 0);
    }
    if (born) memcpy(born, B, sizeof(double) * N);
    if (integrals_out) memcpy(integrals_out, I, sizeof(double) * 2 * N);
    if (dEdI_out) memcpy(dEdI_out, dEdI, sizeof(double) * 2 * N);
    *energy = E;
out:
    grid_free(&c.G);
    free(quad); free(c.pw); free(I); free(B); free(dB); free(dEdB); free(dEdI); free(c.brute_list);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Grid inspection for tests                                                  */
/* ------------------------------------------------------------------------- */

int orc_grid_inspect(const orc_params* p, int N, const double* pos, const double* radius,
                     double* geom, long long* ncell, long long* nentries,
                     int* counts, int* lists, long long counts_cap, long long lists_cap) {
    grid_t G;
    orc_params pp = *p;
    if (pp.policy == ORC_POLICY_BRUTE) pp.policy = ORC_POLICY_GRID;
    int rc = grid_build(&G, &pp, N, pos, radius);
    if (rc) { grid_free(&G); return rc; }
    for (int d = 0; d < 3; ++d) {
        geom[d] = G.min[d];
        geom[3 + d] = G.max[d];
        geom[7 + d] = G.n[d];
    }
    geom[6] = G.g;
    *ncell = G.ncell;
    *nentries = G.start[G.ncell];
    if (counts && counts_cap >= G.ncell)
        for (long long cc = 0; cc < G.ncell; ++cc) counts[cc] = G.start[cc + 1] - G.start[cc];
    if (lists && lists_cap >= G.start[G.ncell]) memcpy(lists, G.atoms, sizeof(int) * (size_t)G.start[G.ncell]);
    grid_free(&G);
    return 0;
}

long long orc_grid_query(const double* geom, const double* point) {
    grid_t G;
    memset(&G, 0, sizeof(G));
    for (int d = 0; d < 3; ++d) {
        G.min[d] = geom[d];
        G.max[d] = geom[3 + d];
        G.n[d] = (int)geom[7 + d];
    }
    G.g = geom[6];
    return grid_cell(&G, point);
}

/* ------------------------------------------------------------------------- */
/* Integrals of any orders and their chain rule (general CharmmGBMVForce      */
/* expressions: CharmmReferenceGBMV.cpp:157-160 evaluates the integrals and   */
/* their gradients, :363-371 applies dE/dI_k(i) dI_k(i)/dx_j)                */
/* ------------------------------------------------------------------------- */

int orc_integral_chain(const orc_params* p, int N, const double* pos, const double* radius, const double* dEdI,
                       double* integrals_out, double* forces_out) {
    if (!p || N <= 0 || !pos || !radius || !integrals_out) return ORC_ERR_ARG;
    if (p->model != ORC_MODEL_GBMV_FORM || p->n_int != 2) return ORC_ERR_ARG;
    if (p->orders[0] < 2 || p->orders[1] < 2) return ORC_ERR_ARG; /* GBSWIntegral.cpp:92-95 */
    if (p->method < ORC_NO_CUTOFF || p->method > ORC_CUTOFF_PERIODIC) return ORC_ERR_ARG;
    if (dEdI && !forces_out) return ORC_ERR_ARG;
    orc_params pp = *p;
    const int type = pp.integral_type;
    ctx_t c;
    memset(&c, 0, sizeof(c));
    c.p = &pp;
    c.N = N;
    c.pos = pos;
    c.R = radius;
    c.nint = 2;
    c.Q = pp.n_radial * pp.n_lebedev;
    const int Q = c.Q;
    double* quad = malloc(sizeof(double) * 5 * Q);
    c.pw = malloc(sizeof(double) * 2 * Q);
    c.brute_list = malloc(sizeof(int) * N);
    int rc = 0;
    if (!quad || !c.pw || !c.brute_list) { rc = ORC_ERR_MEM; goto out; }
    for (int a = 0; a < N; ++a) c.brute_list[a] = a;
    if ((rc = orc_quad_table(pp.n_radial, pp.r0, pp.r1, pp.n_lebedev, quad)) != 0) goto out;
    c.quad = quad;
    for (int k = 0; k < 2; ++k)
        for (int q = 0; q < Q; ++q) c.pw[k * Q + q] = quad[5 * q + 4] / pow(quad[5 * q + 3], pp.orders[k] - 2);
    if ((rc = grid_build(&c.G, &pp, N, pos, radius)) != 0) goto out;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 4) num_threads(pp.nthreads > 1 ? pp.nthreads : 1)
#endif
    for (int i = 0; i < N; ++i) {
        double acc[2] = {0, 0};
        for (int q = 0; q < Q; ++q) {
            double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1], pos[3 * i + 2] + quad[5 * q + 2]};
            const int* list;
            int n = point_list(&c.G, pt, &list, c.brute_list);
            if (type == ORC_INT_GBSW && n == 0) continue;
            gbmv2_sums s2;
            double pre = 0;
            double V = point_volume(&c, pt, list, n, &s2, &pre);
            for (int k = 0; k < 2; ++k) acc[k] += c.pw[k * Q + q] * point_integrand(type, V);
        }
        integrals_out[i] = acc[0];
        integrals_out[N + i] = acc[1];
    }
    if (dEdI) {
        memset(forces_out, 0, sizeof(double) * 3 * N);
        int nt = 1;
#ifdef _OPENMP
        nt = pp.nthreads > 1 ? pp.nthreads : 1;
#endif
        double* part = calloc((size_t)nt * N * 3, sizeof(double));
        if (!part) { rc = ORC_ERR_MEM; goto out; }
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
        {
            int tid = 0;
#ifdef _OPENMP
            tid = omp_get_thread_num();
#endif
            double* g = part + (size_t)tid * N * 3;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
            for (int i = 0; i < N; ++i) {
                for (int q = 0; q < Q; ++q) {
                    double pt[3] = {pos[3 * i] + quad[5 * q], pos[3 * i + 1] + quad[5 * q + 1],
                                    pos[3 * i + 2] + quad[5 * q + 2]};
                    const int* list;
                    int n = point_list(&c.G, pt, &list, c.brute_list);
                    if (type == ORC_INT_GBSW && n == 0) continue;
                    gbmv2_sums s2;
                    double pre = 0;
                    double V = point_volume(&c, pt, list, n, &s2, &pre);
                    double pref = dEdI[i] * c.pw[q] + dEdI[N + i] * c.pw[Q + q];
                    point_gradient(&c, pt, i, list, n, V, &s2, pre, pref, g);
                }
            }
        }
        for (int t = 0; t < nt; ++t)
            for (int a = 0; a < 3 * N; ++a) forces_out[a] -= part[(size_t)t * N * 3 + a];
        free(part);
    }
out:
    grid_free(&c.G);
    free(quad);
    free(c.pw);
    free(c.brute_list);
    return rc;
}
