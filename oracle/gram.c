/* Gram matrices (Gamma) in FP64: direct tetrapyd sum and the Gauss-Legendre
 * factorised form.  CPU restatement, test infrastructure only (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PB 128 /* points per accumulation block */

/* Accumulate G[m][n] (m <= n) += sum_p wq[p][m] * qq[p][n] for np points. */
static void accum_block(int n, int np, const double* wq, const double* qq, double* G)
{
    for (int m = 0; m < n; ++m) {
        double* gm = G + (long)m * n;
        for (int p = 0; p < np; ++p) {
            double c = wq[(long)p * n + m];
            const double* qp = qq + (long)p * n;
            for (int c2 = m; c2 < n; ++c2) gm[c2] += c * qp[c2];
        }
    }
}

static void symmetrise_upper(int n, double* G)
{
    for (int m = 0; m < n; ++m)
        for (int c = m + 1; c < n; ++c) G[(long)c * n + m] = G[(long)m * n + c];
}

/* Gamma_mn = sum over T(lmin,lmax) of permsix * h2 * v1 v2 v3 * Q_m Q_n,
 * walking (i, j, k) as calculate_orthol_3D does (orthogonalisation_late.c:
 * 548-615): i from lmin, j from i, k from j (+1 when i is odd) in steps of 2
 * up to min(i+j, lmax).  Rows i are split statically over threads and the
 * per-thread sums are added in thread order (deterministic). */
void or_gamma_tet(int lmin, int lmax, const double* q, int ldq, const double* v,
                  int n, const int* order, int geom, int i_lo, int i_hi, double* gamma)
{
    int nt = or_max_threads();
    long nn = (long)n * n;
    double* part = (double*)calloc((size_t)nt * nn, sizeof(double));
    int ib = lmin, ie = lmax + 1;
    if (i_hi > 0) { ib = i_lo > lmin ? i_lo : lmin; ie = i_hi < lmax + 1 ? i_hi : lmax + 1; }
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* G = part + (long)t * nn;
        double* wq = (double*)malloc(sizeof(double) * PB * n);
        double* qq = (double*)malloc(sizeof(double) * PB * n);
        int np = 0;
#pragma omp for schedule(static, 1)
        for (int i = ib; i < ie; ++i) {
            for (int j = i; j <= lmax; ++j) {
                int kmax = i + j < lmax ? i + j : lmax;
                for (int k = j + (i & 1); k <= kmax; k += 2) {
                    double h2 = geom ? or_geometric_asym(i, j, k) : or_h2_exact(i, j, k);
                    double w = or_permsix(i, j, k) * h2 * v[i] * v[j] * v[k];
                    for (int m = 0; m < n; ++m) {
                        double Q = or_pijk(q, ldq, order, m, i, j, k);
                        qq[(long)np * n + m] = Q;
                        wq[(long)np * n + m] = w * Q;
                    }
                    if (++np == PB) { accum_block(n, np, wq, qq, G); np = 0; }
                }
            }
        }
        if (np) accum_block(n, np, wq, qq, G);
        free(wq);
        free(qq);
    }
    memset(gamma, 0, sizeof(double) * nn);
    for (int t = 0; t < nt; ++t)
        for (long e = 0; e < nn; ++e) gamma[e] += part[(long)t * nn + e];
    free(part);
    symmetrise_upper(n, gamma);
}

/* Gauss-Legendre factorisation of the same sum (calculate_orthol,
 * orthogonalisation_late.c:239-336; calculate_orthoMij_T 99-129):
 * with int_{-1}^{1} P_l1 P_l2 P_l3 dx = 2 (l1 l2 l3;0 0 0)^2 the cube sum of
 * h2 * f1(l1) f2(l2) f3(l3) equals (1/8pi) int dx prod_a sum_l (2l+1) P_l f_a,
 * exactly on npt >= (3 lmax + 1)/2 nodes; symmetrising the mode products
 * over the cube leaves 1/6 of the 6 pairings of (m_a) with (n_rho(a)). */
void or_gamma_gl(int lmin, int lmax, const double* q, int ldq, const double* v,
                 int n, const int* order, int npt, double* gamma)
{
    int R = 0;
    for (int m = 0; m < n; ++m)
        for (int a = 0; a < 3; ++a)
            if (order[3 * m + a] + 1 > R) R = order[3 * m + a] + 1;
    double* x = (double*)malloc(sizeof(double) * npt);
    double* w = (double*)malloc(sizeof(double) * npt);
    or_gauss_legendre(npt, x, w);
    /* M[i][j][s], full R x R (symmetric) */
    double* M = (double*)calloc((size_t)R * R * npt, sizeof(double));
#pragma omp parallel
    {
        double* P = (double*)malloc(sizeof(double) * (lmax + 1));
        double* acc = (double*)malloc(sizeof(double) * R * R);
#pragma omp for schedule(static)
        for (int s = 0; s < npt; ++s) {
            double xs = x[s];
            P[0] = 1.0;
            if (lmax > 0) P[1] = xs;
            for (int l = 2; l <= lmax; ++l)
                P[l] = ((2.0 * l - 1.0) * xs * P[l - 1] - (l - 1.0) * P[l - 2]) / l;
            memset(acc, 0, sizeof(double) * R * R);
            for (int l = lmin; l <= lmax; ++l) {
                double c = (2.0 * l + 1.0) * v[l] * P[l];
                const double* ql = q + (long)l * ldq;
                for (int i = 0; i < R; ++i) {
                    double ci = c * ql[i];
                    for (int j = i; j < R; ++j) acc[i * R + j] += ci * ql[j];
                }
            }
            for (int i = 0; i < R; ++i)
                for (int j = i; j < R; ++j) {
                    M[((long)i * R + j) * npt + s] = acc[i * R + j];
                    M[((long)j * R + i) * npt + s] = acc[i * R + j];
                }
        }
        free(P);
        free(acc);
    }
    const double scale = 1.0 / (48.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m < n; ++m) {
        int m1 = order[3 * m], m2 = order[3 * m + 1], m3 = order[3 * m + 2];
        for (int c = m; c < n; ++c) {
            int n1 = order[3 * c], n2 = order[3 * c + 1], n3 = order[3 * c + 2];
            const double* A11 = M + ((long)m1 * R + n1) * npt;
            const double* A12 = M + ((long)m1 * R + n2) * npt;
            const double* A13 = M + ((long)m1 * R + n3) * npt;
            const double* A21 = M + ((long)m2 * R + n1) * npt;
            const double* A22 = M + ((long)m2 * R + n2) * npt;
            const double* A23 = M + ((long)m2 * R + n3) * npt;
            const double* A31 = M + ((long)m3 * R + n1) * npt;
            const double* A32 = M + ((long)m3 * R + n2) * npt;
            const double* A33 = M + ((long)m3 * R + n3) * npt;
            double sum = 0.0;
            for (int s = 0; s < npt; ++s) {
                double t = A11[s] * A22[s] * A33[s] + A21[s] * A32[s] * A13[s]
                         + A31[s] * A12[s] * A23[s] + A31[s] * A22[s] * A13[s]
                         + A21[s] * A12[s] * A33[s] + A11[s] * A32[s] * A23[s];
                sum += w[s] * t;
            }
            gamma[(long)m * n + c] = scale * sum;
            gamma[(long)c * n + m] = scale * sum;
        }
    }
    free(M);
    free(x);
    free(w);
}
