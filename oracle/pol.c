/* Polarisation (T/E) parts of the oracle: the mixed XXY walk, the
 * Gauss-Legendre Gram of the four components TTT / TTE / TEE / EEE, and a
 * direct sum over the mixed tetrapyd with the exact h^2 that pins it.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- XXY walk ------------------------------------------------------------ */

/* get_ijk_next_XXY  tetrapyd_tools.c:299-337: i <= j over the X range, k
 * over the Y range from max(j - i, lminY) (parity-bumped) to min(i + j,
 * lmaxY) in steps of 2.  A new i restarts at (i, i, lminY + (2i % 2)), with
 * no parity bump for an odd lminY, as the reference does.  Returns 0 on
 * stepping at or past i + j + k + 1 >= 2 lmaxX + lmaxY, which also sets the
 * last point (lmaxX, lmaxX, lmaxY or lmaxY - 1 by the parity of lmaxY). */
int or_get_ijk_next_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk;
    if (j == lmaxX && k + 2 > lmaxY) {
        ++i;
        j = i;
        k = lminY + (i + j) % 2;
    } else {
        int top = i + j < lmaxY ? i + j : lmaxY;
        if (k + 2 > top) {
            ++j;
            k = j - i > lminY ? j - i : lminY;
            k += (i + j + k) % 2;
        } else {
            k += 2;
        }
    }
    int more = 1;
    if (i + j + k + 1 >= 2 * lmaxX + lmaxY) {
        more = 0;
        i = lmaxX;
        j = lmaxX;
        k = lmaxY % 2 == 0 ? lmaxY : lmaxY - 1;
    }
    *pi = i;
    *pj = j;
    *pk = k;
    return more;
}

/* check_ijk_min_XXY  tetrapyd_tools.c:339-381: clamp up to the minima, order
 * i <= j, then move forward onto the triangle / parity. */
void or_check_ijk_min_XXY(int lminX, int lminY, int* pi, int* pj, int* pk)
{
    int i = *pi < lminX ? lminX : *pi;
    int j = *pj < lminX ? lminX : *pj;
    int k = *pk < lminY ? lminY : *pk;
    if (i > j) { int t = i; i = j; j = t; }
    if (i + j < k) {
        if (i + j < lminY) {
            if (lminY > 2 * lminX) {
                i = lminY / 2;
                j = i + lminY % 2;
                k = lminY;
            } else {
                i = lminX;
                j = lminX;
                k = lminY + lminY % 2;
            }
        } else {
            k = i + j;
        }
    } else if (i + k < j) {
        j = i + k;
    } else if ((i + j + k) % 2 == 1) {
        k += 1;
    }
    *pi = i;
    *pj = j;
    *pk = k;
}

/* check_ijk_max_XXY  tetrapyd_tools.c:383-414: clamp down to the maxima,
 * order i <= j, then move back onto the triangle / parity. */
void or_check_ijk_max_XXY(int lmaxX, int lmaxY, int* pi, int* pj, int* pk)
{
    int i = *pi > lmaxX ? lmaxX : *pi;
    int j = *pj > lmaxX ? lmaxX : *pj;
    int k = *pk > lmaxY ? lmaxY : *pk;
    if (i > j) { int t = i; i = j; j = t; }
    if (i + j < k) k = i + j;
    else if (i + k < j) j = i + k;
    else if ((i + j + k) % 2 == 1) k -= 1;
    *pi = i;
    *pj = j;
    *pk = k;
}

/* the walk's count as decompose_tetrapyd_XXY takes it (tetrapyd_tools.c:
 * 243-251): steps until the walk ends, capped at lmaxX^2 lmaxY, plus one */
long or_count_tetrapyd_XXY(int lminX, int lminY, int lmaxX, int lmaxY)
{
    int i = lminX, j = lminX, k = lminY;
    or_check_ijk_min_XXY(lminX, lminY, &i, &j, &k);
    long limit = (long)lmaxX * lmaxX * lmaxY + 1, count;
    for (count = 1; count < limit; ++count)
        if (!or_get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k)) break;
    return count + 1;
}

long or_walk_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* ijk, long cap)
{
    int i = lminX, j = lminX, k = lminY;
    or_check_ijk_min_XXY(lminX, lminY, &i, &j, &k);
    long n = or_count_tetrapyd_XXY(lminX, lminY, lmaxX, lmaxY);
    for (long p = 0; p < n; ++p) {
        if (p < cap) { ijk[3 * p] = i; ijk[3 * p + 1] = j; ijk[3 * p + 2] = k; }
        if (p + 1 < n) or_get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k);
    }
    return n;
}

/* decompose_tetrapyd_XXY  tetrapyd_tools.c:226-297: split the counted walk
 * over ranks, then threads, walking forward to each slice start. */
void or_decompose_tetrapyd_XXY(or_tetrapyd_limits* block, int rank, int nranks, int nthreads, int lminX,
                               int lminY, int lmaxX, int lmaxY)
{
    int i = lminX, j = lminX, k = lminY;
    or_check_ijk_min_XXY(lminX, lminY, &i, &j, &k);
    long count = or_count_tetrapyd_XXY(lminX, lminY, lmaxX, lmaxY);
    long* per_rank = (long*)malloc(sizeof(long) * nranks);
    long* per_thread = (long*)malloc(sizeof(long) * nthreads);
    or_divide_tasks(count, nranks, per_rank);
    or_divide_tasks(per_rank[rank], nthreads, per_thread);
    for (int r = 0; r < rank; ++r)
        for (long t = 0; t < per_rank[r]; ++t) or_get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k);
    for (int n = 0; n < nthreads; ++n) {
        if (n > 0)
            for (long t = 0; t < per_thread[n - 1]; ++t)
                or_get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k);
        block[n].i_bgn = i;
        block[n].j_bgn = j;
        block[n].k_bgn = k;
        block[n].loops = (int)per_thread[n];
    }
    free(per_rank);
    free(per_thread);
}

/* ---- Gauss-Legendre Gram of the polarisation components ------------------ */

/* M[(i R + j) npt + s] = sum_{l=lmin..lmax} (2l+1) v(l) P_l(x_s) q_i(l) q_j(l)
 * (calculate_orthoMij_T / _E, orthogonalisation_late.c:99-157) */
static double* gl_mtab(int lmin, int lmax, const double* q, int ldq, const double* v, int R, int npt,
                       const double* x)
{
    double* M = (double*)calloc((size_t)R * R * npt, sizeof(double));
#pragma omp parallel
    {
        double* P = (double*)malloc(sizeof(double) * (lmax + 1));
        double* acc = (double*)malloc(sizeof(double) * R * R);
#pragma omp for schedule(static)
        for (int s = 0; s < npt; ++s) {
            P[0] = 1.0;
            if (lmax > 0) P[1] = x[s];
            for (int l = 2; l <= lmax; ++l) P[l] = ((2.0 * l - 1.0) * x[s] * P[l - 1] - (l - 1.0) * P[l - 2]) / l;
            memset(acc, 0, sizeof(double) * R * R);
            for (int l = lmin; l <= lmax; ++l) {
                const double c = (2.0 * l + 1.0) * v[l] * P[l];
                const double* ql = q + (long)l * ldq;
                for (int a = 0; a < R; ++a)
                    for (int b = a; b < R; ++b) acc[a * R + b] += c * ql[a] * ql[b];
            }
            for (int a = 0; a < R; ++a)
                for (int b = a; b < R; ++b) {
                    M[((long)a * R + b) * npt + s] = acc[a * R + b];
                    M[((long)b * R + a) * npt + s] = acc[a * R + b];
                }
        }
        free(P);
        free(acc);
    }
    return M;
}

static int width_of(int n, const int* order, int slot_lo, int slot_hi)
{
    int R = 0;
    for (int m = 0; m < n; ++m)
        for (int a = slot_lo; a <= slot_hi; ++a)
            if (order[3 * m + a] + 1 > R) R = order[3 * m + a] + 1;
    return R;
}

/* calculate_orthol  orthogonalisation_late.c:239-336, one component:
 *   TTT / EEE:  sum_s w_s/(48 pi) sum_{6 perms rho} prod_a M_{m_a n_rho(a)}
 *   TTE:        sum_s w_s/(16 pi) (M^T_{m0 n0} M^T_{m1 n1} + M^T_{m1 n0} M^T_{m0 n1}) M^E_{m2 n2}
 *   TEE:        sum_s w_s/(16 pi) M^T_{m0 n0} (M^E_{m1 n1} M^E_{m2 n2} + M^E_{m2 n1} M^E_{m1 n2})
 * with M^T over [Tlmin, Tlmax] and M^E over [Elmin, Elmax].  order holds
 * the component's modes in find_perm_late_XXX's slot order: TTE = (T, T, E),
 * TEE = (T, E, E).  Output n x n row-major (symmetric). */
void or_gamma_gl_pol(int pol, int Tlmin, int Tlmax, const double* qT, int ldqT, const double* vT, int Elmin,
                     int Elmax, const double* qE, int ldqE, const double* vE, int n, const int* order, int npt,
                     double* gamma)
{
    double* x = (double*)malloc(sizeof(double) * npt);
    double* w = (double*)malloc(sizeof(double) * npt);
    or_gauss_legendre(npt, x, w);
    /* slots of each leg type */
    int tlo = 0, thi = -1, elo = 0, ehi = -1;
    if (pol == 0) thi = 2;
    else if (pol == 1) { thi = 1; elo = 2; ehi = 2; }
    else if (pol == 2) { thi = 0; elo = 1; ehi = 2; }
    else { elo = 0; ehi = 2; }
    const int RT = thi >= tlo ? width_of(n, order, tlo, thi) : 0;
    const int RE = ehi >= elo ? width_of(n, order, elo, ehi) : 0;
    double* MT = RT ? gl_mtab(Tlmin, Tlmax, qT, ldqT, vT, RT, npt, x) : NULL;
    double* ME = RE ? gl_mtab(Elmin, Elmax, qE, ldqE, vE, RE, npt, x) : NULL;
    const double scale = (pol == 1 || pol == 2) ? 1.0 / (16.0 * M_PI) : 1.0 / (48.0 * M_PI);
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m < n; ++m) {
        const int* a = order + 3 * m;
        for (int c = m; c < n; ++c) {
            const int* b = order + 3 * c;
            double sum = 0.0;
            if (pol == 0 || pol == 3) {
                const double* M = pol == 0 ? MT : ME;
                const int R = pol == 0 ? RT : RE;
#define E3(u, v) (M + ((long)a[u] * R + b[v]) * npt)
                const double *A11 = E3(0, 0), *A12 = E3(0, 1), *A13 = E3(0, 2), *A21 = E3(1, 0), *A22 = E3(1, 1),
                             *A23 = E3(1, 2), *A31 = E3(2, 0), *A32 = E3(2, 1), *A33 = E3(2, 2);
#undef E3
                for (int s = 0; s < npt; ++s)
                    sum += w[s] * (A11[s] * A22[s] * A33[s] + A21[s] * A32[s] * A13[s] + A31[s] * A12[s] * A23[s] +
                                   A31[s] * A22[s] * A13[s] + A21[s] * A12[s] * A33[s] + A11[s] * A32[s] * A23[s]);
            } else if (pol == 1) {
                const double* T00 = MT + ((long)a[0] * RT + b[0]) * npt;
                const double* T11 = MT + ((long)a[1] * RT + b[1]) * npt;
                const double* T10 = MT + ((long)a[1] * RT + b[0]) * npt;
                const double* T01 = MT + ((long)a[0] * RT + b[1]) * npt;
                const double* E22 = ME + ((long)a[2] * RE + b[2]) * npt;
                for (int s = 0; s < npt; ++s) sum += w[s] * (T00[s] * T11[s] + T10[s] * T01[s]) * E22[s];
            } else {
                const double* T00 = MT + ((long)a[0] * RT + b[0]) * npt;
                const double* E11 = ME + ((long)a[1] * RE + b[1]) * npt;
                const double* E22 = ME + ((long)a[2] * RE + b[2]) * npt;
                const double* E21 = ME + ((long)a[2] * RE + b[1]) * npt;
                const double* E12 = ME + ((long)a[1] * RE + b[2]) * npt;
                for (int s = 0; s < npt; ++s) sum += w[s] * T00[s] * (E11[s] * E22[s] + E21[s] * E12[s]);
            }
            gamma[(long)m * n + c] = scale * sum;
            gamma[(long)c * n + m] = scale * sum;
        }
    }
    free(MT);
    free(ME);
    free(x);
    free(w);
}

/* ---- direct sum over the mixed tetrapyd (the pin of the above) ----------- */

/* plijk_TTE / plijk_TEE  orthogonalisation_late.c:646-690 at legs (l1, l2,
 * l3), basis rows by multipole */
static double plijk_mixed(int pol, const double* qT, int ldqT, const double* qE, int ldqE, const int* o, int l1,
                          int l2, int l3)
{
    if (pol == 1)
        return 0.5 * (qT[(long)l1 * ldqT + o[0]] * qT[(long)l2 * ldqT + o[1]] +
                      qT[(long)l1 * ldqT + o[1]] * qT[(long)l2 * ldqT + o[0]]) *
               qE[(long)l3 * ldqE + o[2]];
    return 0.5 * qT[(long)l1 * ldqT + o[0]] *
           (qE[(long)l2 * ldqE + o[1]] * qE[(long)l3 * ldqE + o[2]] +
            qE[(long)l2 * ldqE + o[2]] * qE[(long)l3 * ldqE + o[1]]);
}

/* Gamma_mn = sum over the ordered mixed tetrapyd of mult h2 v v v P_m P_n,
 * h2 exact, P = plijk_TTE (pol 1) over i <= j in [Tlmin, Tlmax] (legs 1, 2)
 * and k in [Elmin, Elmax] (leg 3), or P = plijk_TEE (pol 2) over i in the T
 * range (leg 1) and j <= k in the E range (legs 2, 3); mult = 2 for the
 * unequal pair legs, 1 for equal ones (the cube sum over both orders of the
 * pair).  Triangle and even parity by plain loops (not by the walk). */
void or_gamma_xxy_direct(int pol, int Tlmin, int Tlmax, const double* qT, int ldqT, const double* vT, int Elmin,
                         int Elmax, const double* qE, int ldqE, const double* vE, int n, const int* order,
                         double* gamma)
{
    const long nn = (long)n * n;
    /* pair legs over the pair range [plo, phi], single leg over [slo, shi] */
    const int plo = pol == 1 ? Tlmin : Elmin, phi = pol == 1 ? Tlmax : Elmax;
    const int slo = pol == 1 ? Elmin : Tlmin, shi = pol == 1 ? Elmax : Tlmax;
    int nt = or_max_threads();
    double* part = (double*)calloc((size_t)nt * nn, sizeof(double));
#pragma omp parallel num_threads(nt)
    {
        int t = 0;
#ifdef _OPENMP
        t = omp_get_thread_num();
#endif
        double* G = part + (long)t * nn;
        double* P = (double*)malloc(sizeof(double) * n);
#pragma omp for schedule(dynamic, 1)
        for (int a = plo; a <= phi; ++a)
            for (int b = a; b <= phi; ++b)
                for (int c = slo; c <= shi; ++c) {
                    if (c < b - a || c > a + b || (a + b + c) % 2) continue;
                    /* legs: TTE (a, b, c), TEE (c, a, b) */
                    const int l1 = pol == 1 ? a : c, l2 = pol == 1 ? b : a, l3 = pol == 1 ? c : b;
                    const double v1 = vT[l1], v2 = pol == 1 ? vT[l2] : vE[l2], v3 = vE[l3];
                    const double wgt = (a == b ? 1.0 : 2.0) * or_h2_exact(l1, l2, l3) * v1 * v2 * v3;
                    for (int m = 0; m < n; ++m) P[m] = plijk_mixed(pol, qT, ldqT, qE, ldqE, order + 3 * m, l1, l2, l3);
                    for (int m = 0; m < n; ++m) {
                        const double wm = wgt * P[m];
                        for (int c2 = m; c2 < n; ++c2) G[(long)m * n + c2] += wm * P[c2];
                    }
                }
        free(P);
    }
    memset(gamma, 0, sizeof(double) * nn);
    for (int t = 0; t < nt; ++t)
        for (long e = 0; e < nn; ++e) gamma[e] += part[(long)t * nn + e];
    free(part);
    for (int m = 0; m < n; ++m)
        for (int c = m + 1; c < n; ++c) gamma[(long)c * n + m] = gamma[(long)m * n + c];
}

/* ---- polarised cross-Gamma (calculate_gamma.c) --------------------------- */

/* calculate_gamma_2d (spline 0) / calculate_gamma (spline 1), component pol,
 * for all (m, n) (gamma_pt / gamma_2d_pt_XXX calculate_gamma.c:511-771,
 * tables of init_gamma_glint 48-250 and precompute_gammaMij_T / _E 253-352):
 *   M^T[p][q][x](mu_s) = sum_{l=Tlmin..Tlmax} P_l(mu_s) TTw(l) qtT[p][l][x] bT_q(l),
 *     TTw = (2l+1)^(5/6) b_TT / sqrt(b_TT^2 C_TT + N_TT)
 *   M^E[p][q][x](mu_s) = sum_{l=Elmin..Elmax} P_l(mu_s) EEw(l) (s1 qtE - s2 qtT)[p][l][x] bE_q(l),
 *     s1 = b_TT^2 C_TT + N_TT, s2 = sqrt(fsky_E/fsky_T) b_TE^2 C_TE + N_TE,
 *     s3 = b_TT b_EE C_TE + N_TE, s4 = b_EE^2 C_EE + N_EE,
 *     EEw = (2l+1)^(5/6) b_EE / sqrt(s1 (s1 s4 - s3^2))
 *   (both weights 0 for l <= 1);
 *   y_x = x^2 sum_s w_s perm_3x3(M^{X_b}[p_a][q_b][x][s]), X_b = the field of
 *   late slot b (TTT: T T T, TTE: T T E, TEE: T E E, EEE: E E E), zeroed for
 *   x < 10 when |y| > 1e-18; Gamma[m][n] = dphi^2/(8 pi) * spline integral.
 * order_late is the component's table (find_perm_late_XXX slot order); spec
 * holds 9 rows of lsize_q values: C_TT, C_TE, C_EE, N_TT, N_TE, N_EE, b_TT,
 * b_TE, b_EE; qtT / qtE are [np][lsize_q][nx]. */
void or_gamma_cross_gl_pol(int pol, int Tlmin, int Tlmax, const double* bT, int ldbT, int Elmin, int Elmax,
                           const double* bE, int ldbE, const int* order_late, int nlate, int np, int lsize_q,
                           const double* qtT, const double* qtE, int nx, const double* xvec, const int* order_prim,
                           int n, const double* spec, double fsky_ratio, double dphi, int spline, int npt,
                           double* gamma)
{
    static const int fld[4][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const double *CTT = spec, *CTE = spec + lsize_q, *CEE = spec + 2 * lsize_q, *NTT = spec + 3 * lsize_q,
                 *NTE = spec + 4 * lsize_q, *NEE = spec + 5 * lsize_q, *BTT = spec + 6 * lsize_q,
                 *BTE = spec + 7 * lsize_q, *BEE = spec + 8 * lsize_q;
    int R[2] = {0, 0};
    for (int m = 0; m < nlate; ++m)
        for (int b = 0; b < 3; ++b) {
            const int f = fld[pol][b];
            if (order_late[3 * m + b] + 1 > R[f]) R[f] = order_late[3 * m + b] + 1;
        }
    double* gx = (double*)malloc(sizeof(double) * npt);
    double* gw = (double*)malloc(sizeof(double) * npt);
    or_gauss_legendre(npt, gx, gw);
    const int lmaxall = Tlmax > Elmax ? Tlmax : Elmax;
    /* per field: the l weight and the x-resolved primordial side qs[p][l][x] */
    const int lo[2] = {Tlmin, Elmin}, hi[2] = {Tlmax, Elmax}, ldb[2] = {ldbT, ldbE};
    const double* bb[2] = {bT, bE};
    double* M[2] = {NULL, NULL};
    for (int f = 0; f < 2; ++f) {
        if (R[f] == 0) continue;
        double* wl = (double*)calloc((size_t)lmaxall + 1, sizeof(double));
        double* qs = (double*)malloc(sizeof(double) * (size_t)np * lsize_q * nx);
        for (int l = lo[f]; l <= hi[f]; ++l) {
            if (l <= 1) continue;
            const double s1 = BTT[l] * BTT[l] * CTT[l] + NTT[l];
            if (f == 0) {
                wl[l] = pow(2.0 * l + 1.0, 5.0 / 6.0) * (BTT[l] / sqrt(s1));
            } else {
                const double s3 = BTT[l] * BEE[l] * CTE[l] + NTE[l];
                const double s4 = BEE[l] * BEE[l] * CEE[l] + NEE[l];
                wl[l] = pow(2.0 * l + 1.0, 5.0 / 6.0) * (BEE[l] / sqrt(s1 * (s1 * s4 - s3 * s3)));
            }
        }
        for (int p = 0; p < np; ++p)
            for (int l = 0; l < lsize_q; ++l) {
                const double s1 = BTT[l] * BTT[l] * CTT[l] + NTT[l];
                const double s2 = sqrt(fsky_ratio) * BTE[l] * BTE[l] * CTE[l] + NTE[l];
                for (int x = 0; x < nx; ++x) {
                    const long e = ((long)p * lsize_q + l) * nx + x;
                    qs[e] = f == 0 ? qtT[e] : s1 * qtE[e] - s2 * qtT[e];
                }
            }
        /* M[f][((p * R + q) * nx + x) * npt + s] */
        double* Mf = (double*)calloc((size_t)np * R[f] * nx * npt, sizeof(double));
#pragma omp parallel
        {
            double* Pl = (double*)malloc(sizeof(double) * (lmaxall + 1));
#pragma omp for schedule(dynamic, 1)
            for (int s = 0; s < npt; ++s) {
                Pl[0] = 1.0;
                if (lmaxall > 0) Pl[1] = gx[s];
                for (int l = 2; l <= lmaxall; ++l)
                    Pl[l] = ((2.0 * l - 1.0) * gx[s] * Pl[l - 1] - (l - 1.0) * Pl[l - 2]) / l;
                for (int p = 0; p < np; ++p)
                    for (int q = 0; q < R[f]; ++q)
                        for (int x = 0; x < nx; ++x) {
                            double sum = 0.0;
                            for (int l = lo[f]; l <= hi[f]; ++l)
                                sum += Pl[l] * (wl[l] * qs[((long)p * lsize_q + l) * nx + x] * bb[f][(long)l * ldb[f] + q]);
                            Mf[(((long)p * R[f] + q) * nx + x) * npt + s] = sum;
                        }
            }
            free(Pl);
        }
        M[f] = Mf;
        free(wl);
        free(qs);
    }
    const double scale = dphi * dphi / (8.0 * M_PI);
    const int f1 = fld[pol][0], f2 = fld[pol][1], f3 = fld[pol][2];
#pragma omp parallel for schedule(dynamic, 1)
    for (int m = 0; m < nlate; ++m) {
        double* y = (double*)malloc(sizeof(double) * nx);
        const int q1 = order_late[3 * m], q2 = order_late[3 * m + 1], q3 = order_late[3 * m + 2];
        for (int c = 0; c < n; ++c) {
            const int p1 = order_prim[3 * c], p2 = order_prim[3 * c + 1], p3 = order_prim[3 * c + 2];
            for (int x = 0; x < nx; ++x) {
#define MM(f, p, q) (M[f] + (((long)(p) * R[f] + (q)) * nx + x) * npt)
                /* a_ab = M^{X_b}[p_a][q_b] */
                const double *A11 = MM(f1, p1, q1), *A12 = MM(f2, p1, q2), *A13 = MM(f3, p1, q3);
                const double *A21 = MM(f1, p2, q1), *A22 = MM(f2, p2, q2), *A23 = MM(f3, p2, q3);
                const double *A31 = MM(f1, p3, q1), *A32 = MM(f2, p3, q2), *A33 = MM(f3, p3, q3);
#undef MM
                double r = 0.0;
                for (int s = 0; s < npt; ++s)
                    r += gw[s] * (A11[s] * A22[s] * A33[s] + A11[s] * A23[s] * A32[s] + A12[s] * A21[s] * A33[s] +
                                  A12[s] * A23[s] * A31[s] + A13[s] * A21[s] * A32[s] + A13[s] * A22[s] * A31[s]);
                y[x] = xvec[x] * xvec[x] * r;
                if (x < 10 && fabs(y[x]) > 1e-18) y[x] = 0.0;
            }
            const double I = spline ? or_spline_integral_natural(xvec, y, nx) : or_spline_integral_fmm(xvec, y, nx);
            gamma[(long)m * n + c] = I * scale;
        }
        free(y);
    }
    free(M[0]);
    free(M[1]);
    free(gx);
    free(gw);
}
