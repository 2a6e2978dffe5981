/* mcmb_ortho -- orthonormalisation of a Gram file on one GPU, writing the
 * reference's ortho and lambda files (raw n x n FP64).
 *
 *   mcmb_ortho <gamma_file> <n> prim <out_ortho> <out_lambda>
 *     ortho_prim_mpi.c:198-284: ortho = symmetrised D L, lambda = symmetrised
 *     (D L)^-1 (update_ortho / update_lambda), Cholesky of D^-1 G D^-1.
 *   mcmb_ortho <gamma_file> <n> late <out_orthol> <out_lambdal> [pca_flag]
 *     ortho_late_mpi.c:752-842: PCA when the smallest eigenvalue of
 *     D^-1 G D^-1 is negative or pca_flag == 0 (orthol = lambdainv =
 *     D V sqrt(e), lambdal = e^-1/2 V^T D^-1 on the positive eigenvalues),
 *     else the Cholesky factors (lower triangles).
 * Exit status 3 when the Gram is not positive definite in the prim mode
 * (the reference aborts inside GSL there). */
#include "common.h"
#include <string.h>

int main(int argc, char** argv)
{
    if (argc < 6 || (strcmp(argv[3], "prim") && strcmp(argv[3], "late"))) {
        fprintf(stderr, "usage: %s <gamma_file> <n> prim|late <out_ortho> <out_lambda> [pca_flag]\n", argv[0]);
        return 2;
    }
    const int n = atoi(argv[2]);
    const size_t nn = (size_t)n * n;
    double* g = (double*)malloc(sizeof(double) * nn);
    double* o = (double*)malloc(sizeof(double) * nn);
    double* l = (double*)malloc(sizeof(double) * nn);
    double* norm = (double*)malloc(sizeof(double) * n);
    CHECK(mcmb_read_raw(argv[1], (long)nn, g));
    mcmb_ctx* ctx;
    CHECK(mcmb_create(&ctx, 0));
    if (!strcmp(argv[3], "prim")) {
        int bad = -1;
        int st = mcmb_orthonormalise(ctx, n, g, o, l, norm, &bad);
        if (st == MCMB_ERR_NOT_SPD) {
            fprintf(stderr, "not positive definite: %s\n", mcmb_last_error());
            return 3;
        }
        CHECK(st);
        printf("prim orthonormalisation of n=%d\n", n);
    } else {
        const int pca = argc > 6 ? atoi(argv[6]) : 1;
        double* ev = (double*)malloc(sizeof(double) * n);
        int r = 0;
        CHECK(mcmb_orthonormalise_late(ctx, n, g, pca, l, o, norm, ev, &r));
        printf("late orthonormalisation of n=%d: %s, kept %d, eigenvalues %.6e .. %.3e\n", n,
               r < n || pca == 0 ? "PCA" : "Cholesky", r, ev[0], ev[n - 1]);
        free(ev);
    }
    CHECK(mcmb_write_raw(argv[4], (long)nn, o));
    CHECK(mcmb_write_raw(argv[5], (long)nn, l));
    mcmb_destroy(ctx);
    free(g); free(o); free(l); free(norm);
    return 0;
}
