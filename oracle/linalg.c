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

/* Symmetric eigen-decomposition by the cyclic-by-row Jacobi method (Golub &
 * Van Loan, Matrix Computations, 4th ed., Algorithms 8.5.1-8.5.3): rotations
 * (p,q) with t = sign(tau)/(|tau| + sqrt(1 + tau^2)), tau = (a_qq - a_pp)/(2 a_pq),
 * until off(A) <= 1e-15 ||A||_F.  Replaces gsl_eigen_symmv followed by
 * gsl_eigen_symmv_sort(..., GSL_EIGEN_SORT_VAL_DESC) (ortho_prim_mpi.c:223-224,
 * ortho_late_mpi.c:775-776).  evals descending; evecs[i*n + k] = component i
 * of eigenvector k, with its largest-magnitude component made positive (GSL
 * leaves the signs unspecified).  Returns the sweeps used, or -1 if 100
 * sweeps do not converge. */
int or_eigen_sym(int n, const double* A0, double* evals, double* evecs)
{
    double* a = (double*)malloc(sizeof(double) * n * n);
    double* v = (double*)calloc((size_t)n * n, sizeof(double));
    memcpy(a, A0, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) v[(long)i * n + i] = 1.0;
    double fro = 0.0;
    for (long e = 0; e < (long)n * n; ++e) fro += a[e] * a[e];
    int sweep = 0, ok = 0;
    for (; sweep < 100; ++sweep) {
        double off = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                if (i != j) off += a[(long)i * n + j] * a[(long)i * n + j];
        if (off <= 1e-30 * fro) { ok = 1; break; }
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[(long)p * n + q];
                if (apq == 0.0) continue;
                double app = a[(long)p * n + p], aqq = a[(long)q * n + q];
                double tau = (aqq - app) / (2.0 * apq);
                double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
                /* A <- J^T A J, J = [c s; -s c] on (p, q) */
                for (int k = 0; k < n; ++k) {
                    double akp = a[(long)k * n + p], akq = a[(long)k * n + q];
                    a[(long)k * n + p] = c * akp - s * akq;
                    a[(long)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a[(long)p * n + k], aqk = a[(long)q * n + k];
                    a[(long)p * n + k] = c * apk - s * aqk;
                    a[(long)q * n + k] = s * apk + c * aqk;
                }
                a[(long)p * n + p] = app - t * apq;
                a[(long)q * n + q] = aqq + t * apq;
                a[(long)p * n + q] = a[(long)q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vkp = v[(long)k * n + p], vkq = v[(long)k * n + q];
                    v[(long)k * n + p] = c * vkp - s * vkq;
                    v[(long)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    /* sort descending (stable selection by index order for ties) */
    int* idx = (int*)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 1; i < n; ++i) {
        int x = idx[i], j = i - 1;
        while (j >= 0 && a[(long)idx[j] * n + idx[j]] < a[(long)x * n + x]) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = x;
    }
    for (int k = 0; k < n; ++k) {
        int c = idx[k];
        evals[k] = a[(long)c * n + c];
        int imax = 0;
        for (int i = 1; i < n; ++i)
            if (fabs(v[(long)i * n + c]) > fabs(v[(long)imax * n + c])) imax = i;
        double sg = v[(long)imax * n + c] < 0 ? -1.0 : 1.0;
        for (int i = 0; i < n; ++i) evecs[(long)i * n + k] = sg * v[(long)i * n + c];
    }
    free(idx);
    free(v);
    free(a);
    return ok ? sweep : -1;
}

/* Late-side orthonormalisation (ortho_late_mpi.c:752-842) of one Gram:
 * norm_i = sqrt(G_ii), C = D^-1 G D^-1, eigenvalues e (descending) and
 * vectors V of C.  PCA branch when pca_flag == 0 or e_min < 0 (the test at
 * ortho_late_mpi.c:784): r = number of leading positive eigenvalues,
 *   lambda[i][j]    = V[j][i] / (norm_j sqrt(e_i))   for i < r, else 0,
 *   lambdainv[i][j] = norm_i V[i][j] sqrt(e_j)       for j < r, else 0;
 * Cholesky branch otherwise: lambdainv = D L (lower triangle, zero above),
 * lambda = (D L)^-1 (lower).  Returns r (n in the Cholesky branch), -3 if the
 * Cholesky fails or no eigenvalue is positive, -1 if Jacobi does not
 * converge. */
int or_ortho_late(int n, const double* G, int pca_flag, double* lambda, double* lambdainv,
                  double* norm, double* evals)
{
    double* C = (double*)malloc(sizeof(double) * n * n);
    double* V = (double*)malloc(sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) norm[i] = sqrt(G[(long)i * n + i]);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) C[(long)i * n + j] = G[(long)i * n + j] / (norm[i] * norm[j]);
    int st = or_eigen_sym(n, C, evals, V);
    int r = -1;
    if (st < 0) {
        r = -1;
    } else if (pca_flag == 0 || evals[n - 1] < 0.0) {
        r = 0;
        while (r < n && evals[r] > 0.0) ++r;
        if (r == 0) {
            r = -3;
        } else {
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    lambda[(long)i * n + j] = i < r ? V[(long)j * n + i] / (norm[j] * sqrt(evals[i])) : 0.0;
                    lambdainv[(long)i * n + j] = j < r ? norm[i] * V[(long)i * n + j] * sqrt(evals[j]) : 0.0;
                }
        }
    } else {
        double* O = (double*)malloc(sizeof(double) * n * n);
        double* Lm = (double*)malloc(sizeof(double) * n * n);
        int cs = or_orthonormalise(n, G, O, Lm, norm);
        if (cs) {
            r = -3;
        } else {
            for (int i = 0; i < n; ++i)
                for (int j = 0; j < n; ++j) {
                    lambdainv[(long)i * n + j] = j <= i ? O[(long)i * n + j] : 0.0;
                    lambda[(long)i * n + j] = j <= i ? Lm[(long)i * n + j] : 0.0;
                }
            r = n;
        }
        free(O);
        free(Lm);
    }
    free(V);
    free(C);
    return r;
}
