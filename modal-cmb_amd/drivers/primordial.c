/* mcmb_primordial -- reconstruction check of a beta file (primordial_mpi.c:
 * 104-135): prints shape3(k, k, k) of the feature shape against the expansion
 * sum_n beta_n Q_n(k, k, k) along the diagonal, and on a 10 x 10 grid of
 * (k1, k2) at k3 = kmax / 2, with the basis P_p(2 k / kmax - 1).
 *
 *   mcmb_primordial <order_file> <omega> <a2> <eigen_file> [kmax]
 */
#include "common.h"
#include <math.h>

static double expand(const double* beta, const int* order, int n, int pmax, double kmax, double k1, double k2,
                     double k3)
{
    const double ks[3] = {k1, k2, k3};
    double* q = legendre_table(ks, 3, pmax, kmax);
    const int P = pmax + 1;
    double s = 0.0;
    for (int m = 0; m < n; ++m) {
        const int a = order[3 * m], b = order[3 * m + 1], c = order[3 * m + 2];
        const double* q1 = q;
        const double* q2 = q + P;
        const double* q3 = q + 2 * P;
        const double Q = (q1[a] * q2[b] * q3[c] + q1[b] * q2[c] * q3[a] + q1[c] * q2[a] * q3[b] +
                          q1[c] * q2[b] * q3[a] + q1[b] * q2[a] * q3[c] + q1[a] * q2[c] * q3[b]) / 6.0;
        s += beta[m] * Q;
    }
    free(q);
    return s;
}

int main(int argc, char** argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <order_file> <omega> <a2> <eigen_file> [kmax]\n", argv[0]);
        return 2;
    }
    const shape_params sp = {atof(argv[2]), atof(argv[3]), 0.0};
    const double kmax = argc > 5 ? atof(argv[5]) : 0.2;
    int n;
    int* order = read_order(argv[1], &n);
    const int pmax = order_pmax(order, n);
    double* beta = (double*)malloc(sizeof(double) * n);
    CHECK(mcmb_read_raw(argv[4], n, beta));
    printf("# diagonal k1 = k2 = k3\n#   k            shape3         expansion\n");
    double num = 0, d1 = 0, d2 = 0;
    for (int i = 1; i <= 20; ++i) {
        const double k = kmax * i / 20.0;
        const double s = shape3(k, k, k, sp), r = expand(beta, order, n, pmax, kmax, k, k, k);
        printf("%12.6e  %14.7e  %14.7e\n", k, s, r);
    }
    printf("# grid k3 = kmax/2\n#   k1           k2           shape3         expansion\n");
    for (int i = 1; i <= 10; ++i)
        for (int j = 1; j <= 10; ++j) {
            const double k1 = kmax * i / 10.0, k2 = kmax * j / 10.0, k3 = 0.5 * kmax;
            if (k3 > k1 + k2 || k1 > k2 + k3 || k2 > k1 + k3) continue;
            const double s = shape3(k1, k2, k3, sp), r = expand(beta, order, n, pmax, kmax, k1, k2, k3);
            num += s * r; d1 += s * s; d2 += r * r;
            printf("%12.6e %12.6e  %14.7e  %14.7e\n", k1, k2, s, r);
        }
    printf("# cosine over the grid points: %.8f\n", num / sqrt(d1 * d2));
    free(beta); free(order);
    return 0;
}
