/* Shared helpers of the command-line drivers: status checks and the
 * shifted-Legendre basis tables the drivers build. */
#ifndef MCMB_DRIVERS_COMMON_H
#define MCMB_DRIVERS_COMMON_H
#include "modal_cmb.h"
#include <stdio.h>
#include <stdlib.h>

#define CHECK(call)                                                                 \
    do {                                                                            \
        int st_ = (call);                                                           \
        if (st_ != MCMB_OK) {                                                       \
            fprintf(stderr, "%s: %s (%s)\n", #call, mcmb_status_string(st_),       \
                    mcmb_last_error());                                             \
            exit(1);                                                                \
        }                                                                           \
    } while (0)

/* q[i][p] = P_p(2 x_i / xmax - 1) for p <= pmax (basis_functions_legendre,
 * min 0); returns a malloc'd [npts][pmax+1] table. */
static inline double* legendre_table(const double* x, int npts, int pmax, double xmax)
{
    double* q = (double*)malloc(sizeof(double) * (size_t)npts * (pmax + 1));
    double** rows = (double**)malloc(sizeof(double*) * (size_t)npts);
    for (int i = 0; i < npts; ++i) rows[i] = q + (long)i * (pmax + 1);
    basis_functions_legendre(rows, npts, pmax, 0.0, xmax, (double*)x);
    free(rows);
    return q;
}

static inline int order_pmax(const int* order, int n)
{
    int p = 0;
    for (int e = 0; e < 3 * n; ++e) if (order[e] > p) p = order[e];
    return p;
}

/* order table from a file (load_three_int layout) */
static inline int* read_order(const char* path, int* n)
{
    int cap = 1 << 20;
    int* order = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
    CHECK(mcmb_load_order(path, order, cap, n));
    if (*n < 1) { fprintf(stderr, "%s: no modes\n", path); exit(1); }
    return order;
}
#endif
