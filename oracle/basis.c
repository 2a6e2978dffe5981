/* 1D basis, mode order, symmetrised mode products, Gauss-Legendre nodes.
 * CPU restatement, test infrastructure only (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double or_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int or_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* basis_functions.c:88-107: shifted Legendre by the three-term recurrence on
 * x = (v - min)/(max - min), with P1 = 2x - 1 reused in every step. */
void or_basis_legendre(double* a, int size, int pmax, double min, double max, const double* values)
{
    int ld = pmax + 1;
    for (int i = 0; i < size; ++i) {
        double x = (values[i] - min) / (max - min);
        double* row = a + (long)i * ld;
        row[0] = 1.0;
        if (pmax > 0) row[1] = 2.0 * x - 1.0;
        for (int j = 2; j <= pmax; ++j)
            row[j] = ((2.0 * j - 1.0) * row[1] * row[j - 1] - (j - 1.0) * row[j - 2]) / (1.0 * j);
    }
}

/* The explicit order file used by every config: (r,s,t) with r<=s<=t<R in
 * lexicographic order (load_three_int format, io.c:229-258; set_terms_prim
 * case 3, fixed_data.c:172-185). */
int or_order_lex(int R, int* order)
{
    int n = 0;
    for (int r = 0; r < R; ++r)
        for (int s = r; s < R; ++s)
            for (int t = s; t < R; ++t) {
                order[3 * n] = r; order[3 * n + 1] = s; order[3 * n + 2] = t;
                ++n;
            }
    return n;
}

/* orthogonalisation.c:279-295 / orthogonalisation_late.c:620-644 */
double or_pijk(const double* q, int ldq, const int* order, int r, int i, int j, int k)
{
    int p1 = order[3 * r], p2 = order[3 * r + 1], p3 = order[3 * r + 2];
    const double* qi = q + (long)i * ldq;
    const double* qj = q + (long)j * ldq;
    const double* qk = q + (long)k * ldq;
    double b1 = qi[p1] * qj[p2] * qk[p3];
    double b2 = qi[p2] * qj[p3] * qk[p1];
    double b3 = qi[p3] * qj[p1] * qk[p2];
    double b4 = qi[p3] * qj[p2] * qk[p1];
    double b5 = qi[p2] * qj[p1] * qk[p3];
    double b6 = qi[p1] * qj[p3] * qk[p2];
    return (b1 + b2 + b3 + b4 + b5 + b6) / 6.0;
}

/* Gauss-Legendre on [-1,1]: Newton's method on P_n from the Tricomi initial
 * guess, evaluated by the three-term recurrence in long double; weights
 * 2/((1-x^2) P_n'(x)^2).  Nodes ascending. */
void or_gauss_legendre(int n, double* xo, double* wo)
{
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < (n + 1) / 2; ++i) {
        long double th = pi * (4.0L * (i + 1) - 1.0L) / (4.0L * n + 2.0L);
        long double x = cosl(th) * (1.0L - (n - 1.0L) / (8.0L * n * n * n));
        long double dp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = x;
            for (int k = 2; k <= n; ++k) {
                long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                p0 = p1; p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1.0L; }
            dp = n * (x * p1 - p0) / (x * x - 1.0L);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-19L) {
                p0 = 1.0L; p1 = x;
                for (int k = 2; k <= n; ++k) {
                    long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / k;
                    p0 = p1; p1 = p2;
                }
                if (n == 1) { p1 = x; p0 = 1.0L; }
                dp = n * (x * p1 - p0) / (x * x - 1.0L);
                break;
            }
        }
        long double w = 2.0L / ((1.0L - x * x) * dp * dp);
        /* x is the i-th largest node */
        xo[n - 1 - i] = (double)x;
        wo[n - 1 - i] = (double)w;
        xo[i] = (double)(-x);
        wo[i] = (double)w;
    }
    if (n % 2 == 1) xo[n / 2] = 0.0;
}
