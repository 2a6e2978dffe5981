/* Gauss-Legendre quadrature for the library (host side).  Newton's method on
 * P_n from the Tricomi-type starting values x ~ cos(pi (4i+3)/(4n+2)), with P_n
 * and P_n' from the three-term recurrence evaluated in long double; weights
 * 2/((1-x^2) P_n'(x)^2).  Replaces asy() (gl_integration.c:46-168), which
 * needs GSL Bessel functions. */
#include "modal_cmb.h"
#include <math.h>

static void legendre_pn(int n, long double x, long double* pn, long double* dpn)
{
    long double p0 = 1.0L, p1 = x;
    for (int k = 2; k <= n; ++k) {
        long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
        p0 = p1;
        p1 = p2;
    }
    if (n == 0) { p1 = 1.0L; p0 = 0.0L; }
    *pn = p1;
    *dpn = (n == 0) ? 0.0L : n * (x * p1 - p0) / (x * x - 1.0L);
}

/* Roots i0 <= i < i1 of the (n + 1) / 2 in [0, 1) (i counted from -1 up):
 * x[i] = -r, x[n - 1 - i] = r and their weights.  Each root is independent,
 * so ranges may run on separate threads. */
void mcmb_gl_roots(int n, int i0, int i1, double* x, double* w)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int i = i0; i < i1; ++i) {
        long double t = cosl(pi * (4.0L * i + 3.0L) / (4.0L * n + 2.0L));
        long double r = t * (1.0L - (n - 1.0L) / (8.0L * n * (long double)n * n));
        long double pn, dp;
        for (int it = 0; it < 64; ++it) {
            legendre_pn(n, r, &pn, &dp);
            long double step = pn / dp;
            r -= step;
            if (fabsl(step) <= 1e-19L * (1.0L + fabsl(r))) break;
        }
        legendre_pn(n, r, &pn, &dp);
        long double wt = 2.0L / ((1.0L - r * r) * dp * dp);
        x[n - 1 - i] = (double)r;
        x[i] = (double)(-r);
        w[n - 1 - i] = (double)wt;
        w[i] = (double)wt;
        if (i == n - 1 - i) x[i] = 0.0;   /* the middle node of an odd order */
    }
}

int mcmb_gauss_legendre(int n, double* x, double* w)
{
    if (n < 1 || !x || !w) return MCMB_ERR_ARG;
    mcmb_gl_roots(n, 0, (n + 1) / 2, x, w);
    return MCMB_OK;
}
