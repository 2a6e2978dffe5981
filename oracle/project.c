/* Shape templates, the uniform-grid k-space Gram/projection of the reference,
 * and the build's Gauss-Legendre k-tetrapyd Gram/projection.
 * CPU restatement, test infrastructure only (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* calculate_shape.c:933-946: sin(a1 K + 2 * 3.1415927 * a2 / 100)/(k1k2k3)^2,
 * zero when k1k2k3 <= 1e-15 */
double or_feature(double k1, double k2, double k3, double a1, double a2)
{
    double p1 = k1 + k2 + k3, p111 = k1 * k2 * k3;
    double factor = 0.0;
    if (p111 > 1e-15) factor = 1.0 / (p111 * p111);
    return factor * sin(a1 * p1 + 2.0 * 3.1415927 * a2 * 1e-2);
}

/* calculate_shape.c:197-216 */
double or_shape3_feature(double k1, double k2, double k3, double a1, double a2)
{
    double p222 = (k1 * k1) * (k2 * k2) * (k3 * k3);
    return p222 * or_feature(k1, k2, k3, a1, a2);
}

/* orthogonalisation.c:14-86 with the whole prim walk (single-thread
 * semantics; the reference's threaded decomposition integrates block 0 only
 * because its OpenMP pragma is commented out and t is fixed to 0). */
void or_calculate_ortho_uniform(int max, double kmax, const double* q, int ldq,
                                int n, const int* order, double* packed)
{
    double factor = kmax * (1.0 / (double)max);
    double cube = factor * factor * factor;
    or_tetrapyd_limits blk;
    or_decompose_tetrapyd_prim(&blk, 0, 1, 1, 0, max);
    int i = blk.i_bgn, j = blk.j_bgn, k = blk.k_bgn;
    double* b = (double*)malloc(sizeof(double) * n);
    for (int p = 0; p < blk.loops; ++p) {
        double x = or_calculate_weight(0, max, i, j, k) * cube;
        double ks = factor * i + factor * j + factor * k;
        if (ks > 1e-10) ks = 1.0 / ks;
        x *= ks;
        for (int r = 0; r < n; ++r) b[r] = or_pijk(q, ldq, order, r, i, j, k);
        long m = 0;
        for (int r = 0; r < n; ++r)
            for (int s = r; s < n; ++s) packed[m++] += x * b[r] * b[s];
        or_get_ijk_next_prim(max, &i, &j, &k);
    }
    free(b);
}

/* calculate_eigen.c:86-145 for the feature shape (model 4) */
void or_calculate_eigen_uniform(int max, double kmax, const double* q, int ldq, int n,
                                const int* order, double a1, double a2, double* alpha)
{
    double factor = kmax * (1.0 / (double)max);
    double cube = factor * factor * factor;
    or_tetrapyd_limits blk;
    or_decompose_tetrapyd_prim(&blk, 0, 1, 1, 0, max);
    int i = blk.i_bgn, j = blk.j_bgn, k = blk.k_bgn;
    for (int p = 0; p < blk.loops; ++p) {
        double x = or_calculate_weight(0, max, i, j, k) * cube;
        double k1 = factor * i, k2 = factor * j, k3 = factor * k;
        double ks = k1 + k2 + k3;
        if (ks > 1e-10) ks = 1.0 / ks;
        x *= or_shape3_feature(k1, k2, k3, a1, a2) * ks;
        for (int r = 0; r < n; ++r) alpha[r] += x * or_pijk(q, ldq, order, r, i, j, k);
        or_get_ijk_next_prim(max, &i, &j, &k);
    }
}

void or_kgl_limits(int N, const double* kn, int* lim)
{
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double s = kn[i] + kn[j];
            int b = -1;
            for (int c = 0; c < N; ++c)
                if (kn[c] <= s) b = c;
            lim[i * N + j] = b;
        }
}

/* Enumerate the ordered masked GL points a<=b<=c, kn[c] <= kn[a] + kn[b],
 * multiplicity permsix(a,b,c) (nodes are distinct, so index ties are value
 * ties).  Weight mu * w_a w_b w_c / (k_a + k_b + k_c). */
void or_gram_kgl(int N, const double* kn, const double* kw, const double* q, int ldq,
                 int n, const int* order, double* gamma)
{
    int nt = or_max_threads();
    long nn = (long)n * n;
    double* part = (double*)calloc((size_t)nt * nn, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* G = part + (long)t * nn;
        double* Q = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(static, 1)
        for (int a = 0; a < N; ++a)
            for (int b = a; b < N; ++b)
                for (int c = b; c < N; ++c) {
                    if (!(kn[c] <= kn[a] + kn[b])) break;
                    double K = kn[a] + kn[b] + kn[c];
                    double W = or_permsix(a, b, c) * kw[a] * kw[b] * kw[c] / K;
                    for (int m = 0; m < n; ++m) Q[m] = or_pijk(q, ldq, order, m, a, b, c);
                    for (int m = 0; m < n; ++m) {
                        double wm = W * Q[m];
                        double* gm = G + (long)m * n;
                        for (int r = m; r < n; ++r) gm[r] += wm * Q[r];
                    }
                }
        free(Q);
    }
    memset(gamma, 0, sizeof(double) * nn);
    for (int t = 0; t < nt; ++t)
        for (long e = 0; e < nn; ++e) gamma[e] += part[(long)t * nn + e];
    free(part);
    for (int m = 0; m < n; ++m)
        for (int r = m + 1; r < n; ++r) gamma[(long)r * n + m] = gamma[(long)m * n + r];
}

/* alpha_n = sum mu W sin(omega K + phi) Q_n; d/d omega brings K cos(.),
 * d/d phi brings cos(.). */
void or_project_kgl(int N, const double* kn, const double* kw, const double* q, int ldq,
                    int n, const int* order, double omega, double phi,
                    double* alpha, double* d_om, double* d_ph)
{
    int nt = or_max_threads();
    double* part = (double*)calloc((size_t)nt * 3 * n, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* pa = part + (long)t * 3 * n;
#pragma omp for schedule(static, 1)
        for (int a = 0; a < N; ++a)
            for (int b = a; b < N; ++b)
                for (int c = b; c < N; ++c) {
                    if (!(kn[c] <= kn[a] + kn[b])) break;
                    double K = kn[a] + kn[b] + kn[c];
                    double W = or_permsix(a, b, c) * kw[a] * kw[b] * kw[c] / K;
                    double sv = sin(omega * K + phi), cv = cos(omega * K + phi);
                    double f0 = W * sv, f1 = W * K * cv, f2 = W * cv;
                    for (int m = 0; m < n; ++m) {
                        double Q = or_pijk(q, ldq, order, m, a, b, c);
                        pa[m] += f0 * Q;
                        pa[n + m] += f1 * Q;
                        pa[2 * n + m] += f2 * Q;
                    }
                }
    }
    for (int m = 0; m < n; ++m) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int t = 0; t < nt; ++t) {
            s0 += part[(long)t * 3 * n + m];
            s1 += part[(long)t * 3 * n + n + m];
            s2 += part[(long)t * 3 * n + 2 * n + m];
        }
        if (alpha) alpha[m] = s0;
        if (d_om) d_om[m] = s1;
        if (d_ph) d_ph[m] = s2;
    }
    free(part);
}

/* or_project_kgl for nshape parameter pairs theta[b] = (omega, phi) in one
 * pass over the points (the mode products Q_n are formed once per point and
 * shared by the shapes): alpha[b][n], d_om[b][n], d_ph[b][n]. */
void or_project_kgl_batch(int N, const double* kn, const double* kw, const double* q, int ldq,
                          int n, const int* order, int nshape, const double* theta,
                          double* alpha, double* d_om, double* d_ph)
{
    int nt = or_max_threads();
    long per = (long)nshape * 3 * n;
    double* part = (double*)calloc((size_t)nt * per, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* pa = part + (long)t * per;
        double* Q = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 1)
        for (int a = 0; a < N; ++a)
            for (int b = a; b < N; ++b)
                for (int c = b; c < N; ++c) {
                    if (!(kn[c] <= kn[a] + kn[b])) break;
                    double K = kn[a] + kn[b] + kn[c];
                    double W = or_permsix(a, b, c) * kw[a] * kw[b] * kw[c] / K;
                    for (int m = 0; m < n; ++m) Q[m] = or_pijk(q, ldq, order, m, a, b, c);
                    for (int s = 0; s < nshape; ++s) {
                        double om = theta[2 * s], ph = theta[2 * s + 1];
                        double sv = sin(om * K + ph), cv = cos(om * K + ph);
                        double f0 = W * sv, f1 = W * K * cv, f2 = W * cv;
                        double* ps = pa + (long)s * 3 * n;
                        for (int m = 0; m < n; ++m) {
                            ps[m] += f0 * Q[m];
                            ps[n + m] += f1 * Q[m];
                            ps[2 * n + m] += f2 * Q[m];
                        }
                    }
                }
        free(Q);
    }
    for (int s = 0; s < nshape; ++s)
        for (int m = 0; m < n; ++m) {
            double s0 = 0, s1 = 0, s2 = 0;
            for (int t = 0; t < nt; ++t) {
                const double* ps = part + (long)t * per + (long)s * 3 * n;
                s0 += ps[m];
                s1 += ps[n + m];
                s2 += ps[2 * n + m];
            }
            if (alpha) alpha[(long)s * n + m] = s0;
            if (d_om) d_om[(long)s * n + m] = s1;
            if (d_ph) d_ph[(long)s * n + m] = s2;
        }
    free(part);
}
