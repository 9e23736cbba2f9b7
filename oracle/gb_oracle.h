/*
 * gb_oracle.h -- FP64 CPU restatement of the integral Generalized-Born path of
 * ymwangg/OpenMMCustomGBIntegralForce (Reference platform).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by
 * the product library.  Every function cites the reference file:line it follows.
 *
 * Units: nm, kJ/mol, proton charge.
 */
#ifndef GB_ORACLE_H
#define GB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* model */
enum { ORC_MODEL_GBSW_STANDALONE = 0, ORC_MODEL_GBMV_FORM = 1 };
/* CharmmGBMVForce::GBIntegralType (CharmmGBMVForce.h:99-104) */
enum { ORC_INT_GBSW = 0, ORC_INT_GBMV1 = 1, ORC_INT_GBMV2 = 2 };
/* neighbour policy */
enum { ORC_POLICY_GRID = 0,    /* the reference lookup grid of the model */
       ORC_POLICY_BRUTE = 1,   /* every atom, ascending index (reference computeVolume paths) */
       ORC_POLICY_CORNERS = 2 };/* periodic GBMV form: the non-periodic corner rule on the
                                 * periodic lattice instead of the reference's sample lattice
                                 * (only to measure the difference); otherwise GRID */

typedef struct {
    int model;          /* ORC_MODEL_* */
    int integral_type;  /* ORC_INT_* (GBMV-form only; standalone is GBSW) */
    int policy;         /* ORC_POLICY_* */
    int dense;          /* 1: reference loop order with dense dI/dx (N^2) arrays; 0: sparse */
    int nthreads;       /* OpenMP threads for sparse mode (<=1: serial) */
    /* quadrature */
    int n_radial;       /* Gauss-Legendre points on [r0, r1] */
    int n_lebedev;      /* 6, 14, 26, 38 or 50 */
    double r0, r1;
    /* lookup grid */
    double grid_length; /* g */
    double buffer;      /* lookup-table buffer length */
    int cap;            /* atoms per cell, <= 0: uncapped */
    /* integral orders (GBMV-form; standalone uses {4, 7}) */
    int n_int;
    int orders[4];
    /* GBSW smoothing: R' = rscale*(R + roffset), half width sw */
    double sw, rscale, roffset;
    /* GBMV-I / GBMV-II constants */
    double gamma0, beta, lambda, P1, P2, S0;
    double vguard;      /* GBMV-II: |v|^2 below this is treated as S = +inf (V = 1 for beta < 0) */
    /* Born-radius map B = slope / (a0 (1/eta - I4) + a1 (1/(4 eta^4) - I7)^(1/4)) + shift */
    double alpha0, alpha1, eta, slope, shift;
    /* energy prefactor P: E = sum_i P q_i^2 / (2 B_i) + sum_{i<j} P q_i q_j / f_ij */
    double prefactor;
    /* nonbonded method (CharmmGBMVForce.h / CharmmGBSWForce.h NonbondedMethod) */
    int method;         /* ORC_NO_CUTOFF, ORC_CUTOFF_NON_PERIODIC, ORC_CUTOFF_PERIODIC */
    double cutoff;      /* nm */
    double box[3];      /* rectangular periodic box edges (ORC_CUTOFF_PERIODIC) */
    /* excluded pairs of a ParticlePair energy term (CharmmGBMVForce::addExclusion):
     * n_excl pairs of atom indices (2 ints each); GBMV form only */
    int n_excl;
    const int* excl;
} orc_params;

enum { ORC_NO_CUTOFF = 0, ORC_CUTOFF_NON_PERIODIC = 1, ORC_CUTOFF_PERIODIC = 2 };

typedef struct {
    double energy;
    double n_points;     /* quadrature points that found a non-empty neighbour list */
    double n_vol_pairs;  /* sum of list lengths over those points */
    double n_vol_exact;  /* GBSW: point/atom pairs with d < R'+w (grid independent) */
    double n_frc_pairs;  /* pairs with a nonzero chain-rule contribution */
    double n_cells;      /* lookup grid cells */
    double n_entries;    /* sum of cell list lengths */
} orc_stats;

/* Quadrature (CharmmQuadrature.h). */
int orc_gauss_legendre(double x0, double x1, int n, double* x, double* w);
int orc_lebedev(int npts, double* x, double* y, double* z, double* w);
int orc_lebedev_rule_npts(int rule);
/* Q x 5 table (x, y, z, r, w) in the reference order (GBSWIntegral.cpp:41-56). */
int orc_quad_table(int n_radial, double r0, double r1, int n_lebedev, double* out);

/* Default parameters of each model (reference constructors). */
void orc_defaults_gbsw_standalone(orc_params* p);
void orc_defaults_gbmv(orc_params* p, int integral_type);

/*
 * Energy and forces.  pos: N x 3 (row-major), charge / radius: N.
 * energy (scalar) and forces (N x 3) are OVERWRITTEN.  born (N), integrals (n_int x N, integral-major),
 * dEdI (n_int x N) and stats are optional (NULL allowed).
 * Returns 0 on success, < 0 on error.
 */
int orc_compute(const orc_params* p, int N, const double* pos, const double* charge,
                const double* radius, double* energy, double* forces, double* born,
                double* integrals, double* dEdI, orc_stats* stats);

/*
 * General CharmmGBMVForce expressions: the two integrals of orders p->orders[0..1] (any
 * order >= 2; integral-major N x 2) and, when dEdI (2 x N) is given, the chain-rule forces
 * F_j = -sum_i sum_k dEdI_k(i) dI_k(i)/dx_j (N x 3, overwritten).  GBMV form only.
 */
int orc_integral_chain(const orc_params* p, int N, const double* pos, const double* radius, const double* dEdI,
                       double* integrals, double* forces);

/*
 * Lookup-grid inspection (tests): builds the grid of the model and returns its
 * geometry; cells/entries filled when the buffers are large enough.
 * geom: min[3], max[3], g, n[3] (as doubles).  counts: ncell ints, lists: sum(counts).
 */
int orc_grid_inspect(const orc_params* p, int N, const double* pos, const double* radius,
                     double* geom, long long* ncell, long long* nentries,
                     int* counts, int* lists, long long counts_cap, long long lists_cap);
/* Cell of a point (-1 outside) for the grid of the model. */
long long orc_grid_query(const double* geom, const double* point);

#ifdef __cplusplus
}
#endif
#endif
