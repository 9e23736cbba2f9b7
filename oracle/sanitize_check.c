/*
 * sanitize_check.c -- exercises every entry point of the FP64 oracle (gb_oracle.c) on small
 * deterministic systems, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
 * (`make -C oracle sanitize`; tests/test_oracle_sanitized.py runs it).  TEST
 * INFRASTRUCTURE ONLY.  Exits 0 when every call succeeds with finite results and the dense
 * (reference loop order) and sparse forms agree; any memory or undefined-behaviour error
 * aborts the run through the sanitizers.
 */
#include "gb_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NMAX 125

static int fails = 0;

static void check(int ok, const char* what) {
    if (!ok) {
        fprintf(stderr, "sanitize_check: %s\n", what);
        ++fails;
    }
}

/* 5 x 5 x 5 jittered lattice, 0.3 nm apart (a linear congruential generator, seeded). */
static void make_system(int n, double* pos, double* q, double* R, unsigned seed) {
    static const double radii[4] = {0.100, 0.170, 0.185, 0.200};
    unsigned s = seed;
    for (int a = 0; a < n; ++a) {
        const int ix = a % 5, iy = (a / 5) % 5, iz = a / 25;
        const int c[3] = {ix, iy, iz};
        for (int d = 0; d < 3; ++d) {
            s = s * 1664525u + 1013904223u;
            pos[3 * a + d] = 0.3 * c[d] + 0.03 * ((double)(s >> 8) / 16777216.0 - 0.5);
        }
        s = s * 1664525u + 1013904223u;
        q[a] = ((a & 1) ? 0.5 : -0.5) * (0.5 + (double)(s >> 8) / 16777216.0);
        R[a] = radii[(s >> 4) & 3];
    }
}

static int finite_all(const double* v, int n) {
    for (int k = 0; k < n; ++k)
        if (!isfinite(v[k])) return 0;
    return 1;
}

static void run(const orc_params* p0, const char* tag, int n, const double* pos, const double* q, const double* R) {
    double e_sparse = 0.0, e_dense = 0.0, fs[3 * NMAX], fd[3 * NMAX], born[NMAX], integ[2 * NMAX], dedi[2 * NMAX];
    orc_stats st;
    orc_params p = *p0;
    p.dense = 0;
    p.nthreads = 2;
    char msg[160];
    snprintf(msg, sizeof msg, "%s: sparse evaluation failed", tag);
    check(orc_compute(&p, n, pos, q, R, &e_sparse, fs, born, integ, dedi, &st) == 0, msg);
    snprintf(msg, sizeof msg, "%s: non-finite sparse result", tag);
    check(isfinite(e_sparse) && finite_all(fs, 3 * n) && finite_all(born, n), msg);
    p.dense = 1;
    p.nthreads = 1;
    snprintf(msg, sizeof msg, "%s: dense evaluation failed", tag);
    check(orc_compute(&p, n, pos, q, R, &e_dense, fd, NULL, NULL, NULL, NULL) == 0, msg);
    double dmax = 0.0, fbig = 0.0;
    for (int k = 0; k < 3 * n; ++k) {
        dmax = fmax(dmax, fabs(fs[k] - fd[k]));
        fbig = fmax(fbig, fabs(fs[k]));
    }
    snprintf(msg, sizeof msg, "%s: dense and sparse differ (energy %.17g vs %.17g, force %.3g)", tag, e_dense,
             e_sparse, dmax);
    check(fabs(e_dense - e_sparse) <= 1e-10 * fabs(e_sparse) && dmax <= 1e-10 * fbig, msg);
}

int main(void) {
    const int n = NMAX;
    double pos[3 * NMAX], q[NMAX], R[NMAX];
    make_system(n, pos, q, R, 12345u);

    /* quadrature */
    double x[64], w[64], lx[64], ly[64], lz[64], lw[64];
    check(orc_gauss_legendre(0.05, 2.0, 24, x, w) == 0, "gauss-legendre");
    const int rules[5] = {6, 14, 26, 38, 50};
    for (int k = 0; k < 5; ++k) check(orc_lebedev(rules[k], lx, ly, lz, lw) == 0, "lebedev");
    double* table = malloc(sizeof(double) * 5 * 24 * 50);
    check(table && orc_quad_table(24, 0.05, 2.0, 50, table) == 0, "quadrature table");
    free(table);

    /* every model on its reference grid, by brute force, with a cutoff and periodic */
    orc_params p;
    orc_defaults_gbsw_standalone(&p);
    run(&p, "gbsw", n, pos, q, R);
    p.method = ORC_CUTOFF_NON_PERIODIC;
    p.cutoff = 0.8;
    run(&p, "gbsw cutoff", n, pos, q, R);
    p.method = ORC_CUTOFF_PERIODIC;
    p.box[0] = p.box[1] = p.box[2] = 1.6;
    run(&p, "gbsw periodic", n, pos, q, R);
    for (int t = 0; t < 3; ++t) {
        char tag[64];
        orc_defaults_gbmv(&p, t);
        snprintf(tag, sizeof tag, "gbmv type %d", t);
        run(&p, tag, n, pos, q, R);
        p.policy = ORC_POLICY_BRUTE;
        snprintf(tag, sizeof tag, "gbmv type %d brute", t);
        run(&p, tag, 60, pos, q, R);
        orc_defaults_gbmv(&p, t);
        p.method = ORC_CUTOFF_PERIODIC;
        p.cutoff = 0.7;
        p.box[0] = p.box[1] = p.box[2] = 1.6;
        snprintf(tag, sizeof tag, "gbmv type %d periodic", t);
        run(&p, tag, n, pos, q, R);
    }
    /* exclusions of the pair term */
    int excl[8] = {0, 1, 5, 6, 10, 30, 40, 41};
    orc_defaults_gbmv(&p, 2);
    p.n_excl = 4;
    p.excl = excl;
    run(&p, "gbmv exclusions", n, pos, q, R);

    /* the general-expression entry: integrals of orders 3 and 5 and their chain rule */
    orc_defaults_gbmv(&p, 1);
    p.orders[0] = 3;
    p.orders[1] = 5;
    double integ[2 * NMAX], dedi[2 * NMAX], f[3 * NMAX];
    for (int k = 0; k < 2 * n; ++k) dedi[k] = 0.1 * (k % 7) - 0.3;
    check(orc_integral_chain(&p, n, pos, R, dedi, integ, f) == 0 && finite_all(integ, 2 * n) && finite_all(f, 3 * n),
          "integral chain");

    /* grid inspection and queries */
    orc_defaults_gbmv(&p, 2);
    double geom[12];
    long long ncell = 0, nent = 0;
    check(orc_grid_inspect(&p, n, pos, R, geom, &ncell, &nent, NULL, NULL, 0, 0) == 0, "grid size");
    int* counts = malloc(sizeof(int) * (size_t)(ncell > 0 ? ncell : 1));
    int* lists = malloc(sizeof(int) * (size_t)(nent > 0 ? nent : 1));
    check(counts && lists &&
              orc_grid_inspect(&p, n, pos, R, geom, &ncell, &nent, counts, lists, ncell, nent) == 0,
          "grid lists");
    const double pt[3] = {0.6, 0.6, 0.6};
    check(orc_grid_query(geom, pt) >= 0, "grid query");
    free(counts);
    free(lists);

    if (fails) {
        fprintf(stderr, "sanitize_check: %d failure(s)\n", fails);
        return 1;
    }
    printf("sanitize_check: all oracle entry points clean\n");
    return 0;
}
