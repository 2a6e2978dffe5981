/* mcmb_gamma -- the late-time tetrapyd Gram on one GPU (the contraction of
 * gamma_mpi's 3D branch, gamma_mpi.c:316-381, and of calculate_orthol), written
 * in the output_gamma layout (fixed_data.c:3077-3091: raw n x n FP64).
 *
 *   mcmb_gamma <lmin> <lmax> <order_file> <cl> <out_file> [fp64|mixed|gl] [exact|asym]
 *
 * Basis q_p(l) = P_p(2 l / lmax - 1), l = 0..lmax, p up to the largest index
 * of the order table.  Per-leg weight v(l) = 1/C_l with <cl> = "synthetic"
 * (C_l = 1000/(l(l+1))), "modulated" (times 1 + 0.1 sin(l/30)) or a text file
 * of "l C_l" lines ('#' lines skipped) covering lmin..lmax.  The Gram is the
 * unscaled sum over l1 <= l2 <= l3 of permsix h^2 v1 v2 v3 Q_m Q_n. */
#include "common.h"
#include <math.h>
#include <string.h>

static void load_cl(const char* spec, int lmin, int lmax, double* v)
{
    for (int l = 0; l <= lmax; ++l) v[l] = 0.0;
    if (!strcmp(spec, "synthetic") || !strcmp(spec, "modulated")) {
        const int mod = !strcmp(spec, "modulated");
        for (int l = (lmin > 1 ? lmin : 1); l <= lmax; ++l) {
            double cl = 1000.0 / ((double)l * (l + 1));
            if (mod) cl *= 1.0 + 0.1 * sin(l / 30.0);
            v[l] = 1.0 / cl;
        }
        return;
    }
    FILE* fp = fopen(spec, "r");
    if (!fp) { perror(spec); exit(1); }
    char line[256];
    while (fgets(line, sizeof(line), fp)) {
        if (*line == '#') continue;
        double l, cl;
        if (sscanf(line, "%lf %lf", &l, &cl) == 2 && l >= 0 && l <= lmax && cl != 0.0) v[(int)l] = 1.0 / cl;
    }
    fclose(fp);
    for (int l = lmin; l <= lmax; ++l)
        if (v[l] == 0.0) { fprintf(stderr, "%s: no C_l for l = %d\n", spec, l); exit(1); }
}

int main(int argc, char** argv)
{
    if (argc < 6) {
        fprintf(stderr, "usage: %s <lmin> <lmax> <order_file> <synthetic|modulated|cl_file> <out_file> "
                        "[fp64|mixed|gl] [exact|asym]\n", argv[0]);
        return 2;
    }
    const int lmin = atoi(argv[1]), lmax = atoi(argv[2]);
    const char* mode = argc > 6 ? argv[6] : "fp64";
    const int h2 = (argc > 7 && !strcmp(argv[7], "asym")) ? MCMB_H2_ASYMPTOTIC : MCMB_H2_EXACT;
    int n;
    int* order = read_order(argv[3], &n);
    const int pmax = order_pmax(order, n);
    double* lv = (double*)malloc(sizeof(double) * (lmax + 1));
    for (int l = 0; l <= lmax; ++l) lv[l] = l;
    double* q = legendre_table(lv, lmax + 1, pmax, (double)lmax);
    double* v = (double*)malloc(sizeof(double) * (lmax + 1));
    load_cl(argv[4], lmin, lmax, v);
    mcmb_ctx* ctx;
    CHECK(mcmb_create(&ctx, 0));
    CHECK(mcmb_lspace_set(ctx, lmin, lmax, pmax + 1, q, v, n, order, h2));
    double* g = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (!strcmp(mode, "gl")) CHECK(mcmb_gamma_gl(ctx, 0, g, NULL));
    else CHECK(mcmb_gamma_tet(ctx, !strcmp(mode, "mixed") ? MCMB_MIXED : MCMB_FP64, g, NULL));
    double ms;
    int launches;
    CHECK(mcmb_last_kernel_ms(ctx, &ms, &launches));
    printf("gamma %s: lmin=%d lmax=%d modes=%d points=%ld kernel %.3f s -> %s\n", mode, lmin, lmax, n,
           mcmb_lspace_npoints(ctx), ms / 1e3, argv[5]);
    CHECK(mcmb_write_raw(argv[5], (long)n * n, g));
    mcmb_destroy(ctx);
    free(g); free(v); free(q); free(lv); free(order);
    return 0;
}
