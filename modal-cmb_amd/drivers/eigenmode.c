/* mcmb_eigenmode -- projection of the feature shape on the GPU and the beta
 * solve (eigenmode_mpi.c:160-219): alpha_n = <S, Q_n> over the Gauss-Legendre
 * k-tetrapyd, beta = Gamma_k^-1 alpha with the Cholesky factor of the k-space
 * Gram, written in the output_eigen layout (fixed_data.c:2899-2928: raw n FP64).
 *
 *   mcmb_eigenmode <N> <order_file> <omega> <a2> <ortho_file|-> <out_eigen> [kmin kmax]
 *
 * N GL nodes on [kmin, kmax] (default 2e-4, 0.2), basis P_p(2 k / kmax - 1).
 * The shape is sin(omega K + phi)/(k1 k2 k3)^2 with phi = 2 pi a2 / 100 (a2 in
 * percent of 2 pi as calculate_shape.c:933-946 reads it).  ortho_file is the
 * prim-side factor of the k-space Gram (mcmb_ortho ... prim); "-" computes the
 * Gram and its factor here.  When the Gram is not positive definite in FP64
 * (the reference aborts in GSL there), beta is the minimum-norm solution on
 * the directions the late-side PCA keeps: beta = lambda^T lambda alpha. */
#include "common.h"
#include <math.h>
#include <string.h>

int main(int argc, char** argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s <N> <order_file> <omega> <a2> <ortho_file|-> <out_eigen> [kmin kmax]\n", argv[0]);
        return 2;
    }
    const int N = atoi(argv[1]);
    const double omega = atof(argv[3]), a2 = atof(argv[4]);
    const double kmin = argc > 8 ? atof(argv[7]) : 2e-4, kmax = argc > 8 ? atof(argv[8]) : 0.2;
    int n;
    int* order = read_order(argv[2], &n);
    const int pmax = order_pmax(order, n);
    double* x = (double*)malloc(sizeof(double) * N);
    double* w = (double*)malloc(sizeof(double) * N);
    CHECK(mcmb_gauss_legendre(N, x, w));
    for (int a = 0; a < N; ++a) {
        x[a] = 0.5 * (kmax - kmin) * x[a] + 0.5 * (kmax + kmin);
        w[a] = 0.5 * (kmax - kmin) * w[a];
    }
    double* q = legendre_table(x, N, pmax, kmax);
    mcmb_ctx* ctx;
    CHECK(mcmb_create(&ctx, 0));
    CHECK(mcmb_kspace_set(ctx, N, x, w, pmax + 1, q, n, order));
    const double theta[2] = {omega, 2.0 * 3.1415927 * a2 * 1e-2};
    double* alpha = (double*)malloc(sizeof(double) * n);
    double* beta = (double*)malloc(sizeof(double) * n);
    CHECK(mcmb_project_fwd(ctx, MCMB_FP64, 1, theta, alpha, NULL));
    const size_t nn = (size_t)n * n;
    double* ortho = (double*)malloc(sizeof(double) * nn);
    double* lam = (double*)malloc(sizeof(double) * nn);
    double* norm = (double*)malloc(sizeof(double) * n);
    double* g = NULL;
    int spd = 1;
    if (!strcmp(argv[5], "-")) {
        g = (double*)malloc(sizeof(double) * nn);
        CHECK(mcmb_kgram(ctx, MCMB_FP64, g, NULL));
        int bad = -1;
        int st = mcmb_orthonormalise(ctx, n, g, ortho, lam, norm, &bad);
        if (st == MCMB_ERR_NOT_SPD) spd = 0;
        else CHECK(st);
    } else {
        CHECK(mcmb_read_raw(argv[5], (long)nn, ortho));
    }
    if (spd) {
        CHECK(mcmb_cholesky_solve(ctx, n, ortho, 1, alpha, beta));
        printf("beta by Cholesky solve, n=%d\n", n);
    } else {
        double* linv = (double*)malloc(sizeof(double) * nn);
        double* ev = (double*)malloc(sizeof(double) * n);
        int r = 0;
        CHECK(mcmb_orthonormalise_late(ctx, n, g, 0, lam, linv, norm, ev, &r));
        double* y = (double*)calloc(n, sizeof(double));
        for (int i = 0; i < r; ++i)
            for (int j = 0; j < n; ++j) y[i] += lam[(long)i * n + j] * alpha[j];
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int i = 0; i < r; ++i) s += lam[(long)i * n + j] * y[i];
            beta[j] = s;
        }
        printf("k-space Gram not positive definite in FP64 (%s); beta on the %d kept PCA directions of %d\n",
               mcmb_last_error(), r, n);
        free(y); free(ev); free(linv);
    }
    double sa = 0, sb = 0;
    for (int m = 0; m < n; ++m) { sa += alpha[m] * alpha[m]; sb += beta[m] * beta[m]; }
    printf("omega=%g a2=%g: |alpha|=%.6e |beta|=%.6e -> %s\n", omega, a2, sqrt(sa), sqrt(sb), argv[6]);
    CHECK(mcmb_write_raw(argv[6], n, beta));
    mcmb_destroy(ctx);
    free(g); free(norm); free(lam); free(ortho); free(beta); free(alpha); free(q); free(w); free(x); free(order);
    return 0;
}
