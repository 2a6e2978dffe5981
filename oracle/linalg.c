/* FP64 orthonormalisation and solves (ortho_prim_mpi.c:198-284,
 * eigenmode_mpi.c:195-219).  CPU restatement, test infrastructure only. */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* In-place lower Cholesky of a row-major n x n SPD matrix (lower triangle
 * written, upper untouched).  Returns 0 or -3 when a pivot is not positive. */
static int chol_lower(int n, double* A)
{
    for (int j = 0; j < n; ++j) {
        double d = A[(long)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(long)j * n + k] * A[(long)j * n + k];
        if (!(d > 0.0)) return -3;
        d = sqrt(d);
        A[(long)j * n + j] = d;
#pragma omp parallel for schedule(static)
        for (int i = j + 1; i < n; ++i) {
            double s = A[(long)i * n + j];
            const double* ai = A + (long)i * n;
            const double* aj = A + (long)j * n;
            for (int k = 0; k < j; ++k) s -= ai[k] * aj[k];
            A[(long)i * n + j] = s / d;
        }
    }
    return 0;
}

/* ortho_prim_mpi.c:198-284: normalise by sqrt(diag), Cholesky C = L L^T,
 * ortho[i][j] = ortho[j][i] = norm[max] L[max][min] (symmetrised D L as
 * update_ortho leaves it), lambda = (D L)^-1 symmetrised the same way. */
int or_orthonormalise(int n, const double* G, double* ortho, double* lambda, double* norm)
{
    double* L = (double*)malloc(sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) norm[i] = sqrt(G[(long)i * n + i]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) L[(long)i * n + j] = G[(long)i * n + j] / (norm[i] * norm[j]);
    int st = chol_lower(n, L);
    if (st) { free(L); return st; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double x = norm[i] * L[(long)i * n + j];
            ortho[(long)i * n + j] = x;
            ortho[(long)j * n + i] = x;
        }
    /* Linv = L^-1 (lower) column by column by forward substitution */
    double* Li = (double*)calloc((size_t)n * n, sizeof(double));
#pragma omp parallel for schedule(dynamic, 4)
    for (int c = 0; c < n; ++c) {
        Li[(long)c * n + c] = 1.0 / L[(long)c * n + c];
        for (int i = c + 1; i < n; ++i) {
            double s = 0.0;
            for (int k = c; k < i; ++k) s -= L[(long)i * n + k] * Li[(long)k * n + c];
            Li[(long)i * n + c] = s / L[(long)i * n + i];
        }
    }
    /* (D L)^-1 = L^-1 D^-1 : entry (i, j) = Linv(i, j) / norm[j], i >= j */
    for (int i = 0; i < n; ++i)
        for (int j = 0; j <= i; ++j) {
            double x = Li[(long)i * n + j] / norm[j];
            lambda[(long)i * n + j] = x;
            lambda[(long)j * n + i] = x;
        }
    free(Li);
    free(L);
    return 0;
}

/* beta = Gamma^-1 alpha with Gamma = F F^T, F = the lower half of the
 * symmetrised factor (gsl_linalg_cholesky_solve, eigenmode_mpi.c:211). */
void or_cholesky_solve(int n, const double* F, const double* alpha, double* beta)
{
    double* y = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) {
        double s = alpha[i];
        for (int k = 0; k < i; ++k) s -= F[(long)i * n + k] * y[k];
        y[i] = s / F[(long)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < n; ++k) s -= F[(long)k * n + i] * beta[k];
        beta[i] = s / F[(long)i * n + i];
    }
    free(y);
}
