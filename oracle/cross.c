/* Cross-Gamma of the late-time basis against the projected primordial basis
 * q-tilde: the 3D tetrapyd form with the trapezium rule in x (gamma_3d_host)
 * and the Gauss-Legendre form with a cubic spline in x (calculate_gamma_2d /
 * calculate_gamma).  CPU restatement, test infrastructure only (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* init_gamma_3Dint, calculate_gamma_3d.c:23-61 (TT part): lweight3D[l] =
 * (2l+1)^(-1/3) and q3[l][p][x] = (2l+1)^(1/6) / sqrt(C_l + N_l / b_l^2)
 * qt[p][l][x] for l > 1, both 0 for l <= 1.  The l index of the q-tilde
 * table is the multipole (lvec_T[l] = l). */
void or_qtilde3d(int np, int lsize, int nx, const double* qt, const double* cl, const double* nl,
                 const double* bl, double* q3, double* lw)
{
    for (int l = 0; l < lsize; ++l) {
        double s4 = 0.0;
        lw[l] = 0.0;
        if (l > 1) {
            lw[l] = 1.0 / pow(2.0 * l + 1.0, 1.0 / 3.0);
            double s1 = pow(2.0 * l + 1.0, 1.0 / 6.0);
            double s2 = cl[l] + nl[l] / (bl[l] * bl[l]);
            s4 = s1 * (1.0 / sqrt(s2));
        }
        for (int x = 0; x < nx; ++x)
            for (int p = 0; p < np; ++p)
                q3[((long)l * np + p) * nx + x] = s4 * qt[((long)p * lsize + l) * nx + x];
    }
}

/* calculate_sijk_TTT with integrate() under TRAPEZIUM, calculate_gamma_3d.
 * c:783-860: y_x = x^2 (1/6) sum of the 6 slot permutations of the products
 * q3[l1][a_1] q3[l2][a_2] q3[l3][a_3], zeroed for x < 10 when |y_x| > 1e-18,
 * then sum_x (x_{x+1} - x_x) (y_x + y_{x+1}) / 2. */
static double sijk(const double* q3, int np, int nx, const double* xv, const double* xd, int a1, int a2, int a3,
                   int l1, int l2, int l3, double* y)
{
    const double* Q1 = q3 + (long)l1 * np * nx;
    const double* Q2 = q3 + (long)l2 * np * nx;
    const double* Q3 = q3 + (long)l3 * np * nx;
    for (int i = 0; i < nx; ++i) {
        double p1 = Q1[a1 * nx + i], p2 = Q1[a2 * nx + i], p3 = Q1[a3 * nx + i];
        double p4 = Q2[a1 * nx + i], p5 = Q2[a2 * nx + i], p6 = Q2[a3 * nx + i];
        double p7 = Q3[a1 * nx + i], p8 = Q3[a2 * nx + i], p9 = Q3[a3 * nx + i];
        double t = p1 * p5 * p9 + p2 * p6 * p7 + p3 * p4 * p8 + p3 * p5 * p7 + p2 * p4 * p9 + p1 * p6 * p8;
        y[i] = xv[i] * xv[i] * t * (1.0 / 6.0);
    }
    for (int i = 0; i < nx && i < 10; ++i)
        if (fabs(y[i]) > 1e-18) y[i] = 0.0;
    double r = 0.0;
    for (int i = 0; i < nx - 1; ++i) r += xd[i] * (y[i] + y[i + 1]) * 0.5;
    return r;
}

/* calculate_gamma_TTT_3D_nested, calculate_gamma_3d.c:274-440, whole walk
 * (one rank): over T(2, lsize-1) in the loop nest's order, z = permsix h2
 * lw(l1) lw(l2) lw(l3) (h2 asymptotic, calculate_geometric, when geom = 1;
 * exact when geom = 0), Gamma[m][n] += plijk_TTT(m, l) z x[l][n] with the
 * late basis b[l][ldb] (row = multipole, plijk 0 outside [tlmin, tlmax],
 * orthogonalisation_late.c:620-644) and x from sijk; scaled by 6 dphi^2 at
 * the end (437).  Gamma is n x n, rows late modes, columns primordial modes.
 * Point range [p0, p1) of the walk only when p1 > 0. */
void or_gamma_cross_3d(int lsize, int tlmin, int tlmax, const double* b, int ldb, const int* order_late, int nlate,
                       int np, const double* q3, int nx, const double* xvec, const int* order_prim, int n,
                       const double* lw, int geom, double dphi, long p0, long p1, double* gamma)
{
    double* xd = (double*)malloc(sizeof(double) * nx);
    for (int i = 0; i < nx - 1; ++i) xd[i] = xvec[i + 1] - xvec[i];
    /* points of the walk in order */
    long npts = 0;
    for (int i = 2; i < lsize; ++i)
        for (int j = i; j < lsize; ++j)
            for (int k = j + (i % 2); k <= (i + j < lsize - 1 ? i + j : lsize - 1); k += 2) ++npts;
    int* P = (int*)malloc(sizeof(int) * 3 * (size_t)npts);
    long e = 0;
    for (int i = 2; i < lsize; ++i)
        for (int j = i; j < lsize; ++j)
            for (int k = j + (i % 2); k <= (i + j < lsize - 1 ? i + j : lsize - 1); k += 2) {
                P[3 * e] = i; P[3 * e + 1] = j; P[3 * e + 2] = k; ++e;
            }
    long pb = 0, pe = npts;
    if (p1 > 0) { pb = p0; pe = p1 < npts ? p1 : npts; }
    int nt = or_max_threads();
    long nn = (long)nlate * n;
    double* part = (double*)calloc((size_t)nt * nn, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* G = part + (long)t * nn;
        double* y = (double*)malloc(sizeof(double) * nx);
        double* xs = (double*)malloc(sizeof(double) * n);
        double* pz = (double*)malloc(sizeof(double) * nlate);
#pragma omp for schedule(dynamic, 64)
        for (long q = pb; q < pe; ++q) {
            int i = P[3 * q], j = P[3 * q + 1], k = P[3 * q + 2];
            int inr = i >= tlmin && i <= tlmax && j >= tlmin && j <= tlmax && k >= tlmin && k <= tlmax;
            if (!inr) continue;   /* plijk_TTT = 0 */
            double h2 = geom ? or_geometric_asym(i, j, k) : or_h2_exact(i, j, k);
            double z = or_permsix(i, j, k) * h2 * lw[i] * lw[j] * lw[k];
            for (int c = 0; c < n; ++c)
                xs[c] = sijk(q3, np, nx, xvec, xd, order_prim[3 * c], order_prim[3 * c + 1], order_prim[3 * c + 2],
                             i, j, k, y);
            for (int m = 0; m < nlate; ++m) pz[m] = or_pijk(b, ldb, order_late, m, i, j, k) * z;
            for (int m = 0; m < nlate; ++m)
                for (int c = 0; c < n; ++c) G[(long)m * n + c] += pz[m] * xs[c];
        }
        free(y);
        free(xs);
        free(pz);
    }
    memset(gamma, 0, sizeof(double) * nn);
    for (int t = 0; t < nt; ++t)
        for (long q = 0; q < nn; ++q) gamma[q] += part[(long)t * nn + q];
    for (long q = 0; q < nn; ++q) gamma[q] *= 6.0 * dphi * dphi;
    free(part);
    free(P);
    free(xd);
}

/* integrate_cubic_spline, calculate_gamma.c:845-934: the spline whose end
 * third derivatives match the cubics through the first and last four points
 * (Forsythe, Malcolm and Moler), integrated piece by piece from its
 * coefficients. */
double or_spline_integral_fmm(const double* x, const double* y, int n)
{
    if (n < 2) return 0.0;
    if (n == 3) return NAN;   /* the end conditions read a fourth point (the reference reads past its arrays) */
    double *b = (double*)malloc(sizeof(double) * n), *c = (double*)malloc(sizeof(double) * n),
           *d = (double*)malloc(sizeof(double) * n);
    double *sx = (double*)malloc(sizeof(double) * n), *sy = (double*)malloc(sizeof(double) * n);
    for (int i = 0; i < n; ++i) { sx[i] = x[i]; sy[i] = y[i]; }
    /* spl[2] = b (diagonal, then linear coefficient), spl[3] = c, spl[4] = d */
    d[0] = sx[1] - sx[0];
    c[1] = (sy[1] - sy[0]) / d[0];
    for (int i = 1; i < n - 1; ++i) {
        d[i] = x[i + 1] - sx[i];
        b[i] = 2.0 * (d[i - 1] + d[i]);
        c[i + 1] = (sy[i + 1] - sy[i]) / d[i];
        c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[n - 1] = -d[n - 2];
    c[0] = 0.0;
    c[n - 1] = 0.0;
    if (n > 2) {
        c[0] = c[2] / (sx[3] - sx[1]) - c[1] / (sx[2] - sx[0]);
        c[n - 1] = c[n - 2] / (sx[n - 1] - sx[n - 3]) - c[n - 3] / (sx[n - 2] - sx[n - 4]);
        c[0] = c[0] * d[0] * d[0] / (sx[3] - sx[0]);
        c[n - 1] = -c[n - 1] * d[n - 2] * d[n - 2] / (sx[n - 1] - sx[n - 4]);
    }
    for (int i = 1; i < n; ++i) {
        double h = d[i - 1] / b[i - 1];
        b[i] = b[i] - h * d[i - 1];
        c[i] = c[i] - h * c[i - 1];
    }
    c[n - 1] = c[n - 1] / b[n - 1];
    for (int i = n - 2; i >= 0; --i) c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    b[n - 1] = (sy[n - 1] - sy[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2.0 * c[n - 1]);
    for (int i = 0; i < n - 1; ++i) {
        b[i] = (sy[i + 1] - sy[i]) / d[i] - d[i] * (c[i + 1] + 2.0 * c[i]);
        d[i] = (c[i + 1] - c[i]) / d[i];
        c[i] = 3.0 * c[i];
    }
    c[n - 1] = 3.0 * c[n - 1];
    d[n - 1] = d[n - 2];
    double s = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        double h = x[i + 1] - x[i];
        s += h * (sy[i] + h * (b[i] / 2.0 + h * (c[i] / 3.0 + h * d[i] / 4.0)));
    }
    free(b); free(c); free(d); free(sx); free(sy);
    return s;
}

/* gsl_spline_eval_integ of a gsl_interp_cspline over [x_0, x_{n-1}]
 * (calculate_gamma.c:985-998): the natural cubic spline (second derivative 0
 * at both ends); integral sum_i h_i (y_i + y_{i+1}) / 2 - h_i^3 (M_i +
 * M_{i+1}) / 24 with the knot second derivatives M from the tridiagonal
 * system (Thomas algorithm). */
double or_spline_integral_natural(const double* x, const double* y, int n)
{
    if (n < 2) return 0.0;
    double* M = (double*)calloc((size_t)n, sizeof(double));
    if (n > 2) {
        int m = n - 2;
        double *a = (double*)malloc(sizeof(double) * m), *bd = (double*)malloc(sizeof(double) * m),
               *cc = (double*)malloc(sizeof(double) * m), *r = (double*)malloc(sizeof(double) * m);
        for (int i = 1; i <= m; ++i) {
            double h0 = x[i] - x[i - 1], h1 = x[i + 1] - x[i];
            a[i - 1] = h0;
            bd[i - 1] = 2.0 * (h0 + h1);
            cc[i - 1] = h1;
            r[i - 1] = 6.0 * ((y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0);
        }
        for (int i = 1; i < m; ++i) {
            double w = a[i] / bd[i - 1];
            bd[i] -= w * cc[i - 1];
            r[i] -= w * r[i - 1];
        }
        M[m] = r[m - 1] / bd[m - 1];
        for (int i = m - 2; i >= 0; --i) M[i + 1] = (r[i] - cc[i] * M[i + 2]) / bd[i];
        free(a); free(bd); free(cc); free(r);
    }
    double s = 0.0;
    for (int i = 0; i < n - 1; ++i) {
        double h = x[i + 1] - x[i];
        s += h * (y[i] + y[i + 1]) / 2.0 - h * h * h * (M[i] + M[i + 1]) / 24.0;
    }
    free(M);
    return s;
}

/* init_gamma_glint (calculate_gamma.c:48-240, TT part), precompute_gammaMij_T
 * (255-303), gamma_2d_pt_TTT (511-555) and calculate_gamma_2d (772-843) /
 * calculate_gamma (936-1016):
 *   TTweight(l) = (2l+1)^(5/6) b_l / sqrt(b_l^2 C_l + N_l) (0 for l <= 1)
 *   M[p][q][x][s] = sum_{l=tlmin..tlmax} P_l(mu_s) TTweight(l) qt[p][l][x] b_q(l)
 *   y_x = x^2 sum_s w_s perm_3x3(M[p_a][q_b][x][s]), zeroed for x < 10 when
 *         |y_x| > 1e-18
 *   Gamma[m][n] = dphi^2 / (8 pi) * spline integral of y over x
 * with (p_a) = order_prim[n], (q_b) = order_late[m]; spline 0: the FMM end
 * conditions of integrate_cubic_spline (calculate_gamma_2d), 1: the natural
 * spline of gsl_interp_cspline (calculate_gamma).  npt Gauss-Legendre nodes
 * (the reference: 3 lmax / 2 + 1 from asy(); here the textbook rule). */
void or_gamma_cross_gl(int tlmin, int tlmax, const double* b, int ldb, const int* order_late, int nlate, int np,
                       int lsize_q, const double* qt, int nx, const double* xvec, const int* order_prim, int n,
                       const double* cl, const double* nl, const double* bl, double dphi, int spline, int npt,
                       double* gamma)
{
    int Rq = 0;
    for (int m = 0; m < 3 * nlate; ++m) if (order_late[m] + 1 > Rq) Rq = order_late[m] + 1;
    double* gx = (double*)malloc(sizeof(double) * npt);
    double* gw = (double*)malloc(sizeof(double) * npt);
    or_gauss_legendre(npt, gx, gw);
    double* tw = (double*)calloc((size_t)tlmax + 1, sizeof(double));
    for (int l = tlmin; l <= tlmax; ++l)
        if (l > 1) {
            double s1 = pow(2.0 * l + 1.0, 5.0 / 6.0);
            double s2 = bl[l] * bl[l] * cl[l] + nl[l];
            tw[l] = s1 * (bl[l] / sqrt(s2));
        }
    /* M[((p * Rq + q) * nx + x) * npt + s] */
    double* M = (double*)calloc((size_t)np * Rq * nx * npt, sizeof(double));
#pragma omp parallel
    {
        double* Pl = (double*)malloc(sizeof(double) * (tlmax + 1));
        double* vec = (double*)malloc(sizeof(double) * (tlmax + 1));
#pragma omp for schedule(dynamic, 1)
        for (int s = 0; s < npt; ++s) {
            double xs = gx[s];
            Pl[0] = 1.0;
            if (tlmax > 0) Pl[1] = xs;
            for (int l = 2; l <= tlmax; ++l) Pl[l] = ((2.0 * l - 1.0) * xs * Pl[l - 1] - (l - 1.0) * Pl[l - 2]) / l;
            for (int p = 0; p < np; ++p)
                for (int q = 0; q < Rq; ++q)
                    for (int x = 0; x < nx; ++x) {
                        double sum = 0.0;
                        for (int l = tlmin; l <= tlmax; ++l)
                            sum += Pl[l] * (tw[l] * qt[((long)p * lsize_q + l) * nx + x] * b[(long)l * ldb + q]);
                        M[(((long)p * Rq + q) * nx + x) * npt + s] = sum;
                    }
        }
        free(Pl);
        free(vec);
    }
    const double scale = dphi * dphi / (8.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m < nlate; ++m) {
        double* y = (double*)malloc(sizeof(double) * nx);
        const int q1 = order_late[3 * m], q2 = order_late[3 * m + 1], q3 = order_late[3 * m + 2];
        for (int c = 0; c < n; ++c) {
            const int p1 = order_prim[3 * c], p2 = order_prim[3 * c + 1], p3 = order_prim[3 * c + 2];
            for (int x = 0; x < nx; ++x) {
#define MM(p, q) (M + (((long)(p) * Rq + (q)) * nx + x) * npt)
                const double *A1 = MM(p1, q1), *A2 = MM(p1, q2), *A3 = MM(p1, q3);
                const double *A4 = MM(p2, q1), *A5 = MM(p2, q2), *A6 = MM(p2, q3);
                const double *A7 = MM(p3, q1), *A8 = MM(p3, q2), *A9 = MM(p3, q3);
#undef MM
                double r = 0.0;
                for (int s = 0; s < npt; ++s) {
                    double t = A1[s] * A5[s] * A9[s] + A1[s] * A6[s] * A8[s] + A2[s] * A4[s] * A9[s]
                             + A2[s] * A6[s] * A7[s] + A3[s] * A4[s] * A8[s] + A3[s] * A5[s] * A7[s];
                    r += gw[s] * t;
                }
                y[x] = xvec[x] * xvec[x] * r;
                if (x < 10 && fabs(y[x]) > 1e-18) y[x] = 0.0;
            }
            double I = spline ? or_spline_integral_natural(xvec, y, nx) : or_spline_integral_fmm(xvec, y, nx);
            gamma[(long)m * n + c] = I * scale;
        }
        free(y);
    }
    free(M);
    free(tw);
    free(gx);
    free(gw);
}
