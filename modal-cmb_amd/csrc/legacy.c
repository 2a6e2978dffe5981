/* Reference-named entry points of PetarSuman/Modal-CMB, implemented over the
 * library's own host code and the bound GPU context.  Same names, argument
 * meaning and array layouts as the reference; errors are recorded (query
 * mcmb_legacy_status / mcmb_last_error) instead of exiting. */
#include "modal_cmb.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* The caches of the per-(m, n) entry points are filled on first use under
 * g_mu, so those entry points may be called from several threads at once
 * (calculate_gamma_2d runs inside an omp for, gamma_mpi.c:672-688). */
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static mcmb_ctx* g_ctx = NULL;
static int g_lsize = 0;                 /* lmax + 1 of the bound l-space setup */
static _Thread_local int g_status = MCMB_OK;   /* last status of this thread's void legacy call */
static double* g_gamma3d = NULL;        /* cached FP64 tetrapyd Gram, n x n */
static double* g_gammagl = NULL;        /* cached GL-factorised Gram, n x n */
static double* g_cross[2] = {NULL, NULL};   /* cached GL cross-Gamma, FMM / natural spline */
static long g_cross_gen[2] = {-1, -1};  /* q-tilde generation of each cross cache */
static double* g_pol[4] = {NULL, NULL, NULL, NULL};   /* cached polarised GL Grams, TTT..EEE */
static double* g_crossp[2][4];          /* cached polarised GL cross-Gamma, per x rule and component */
static long g_crossp_gen[2][2] = {{-1, -1}, {-1, -1}};   /* (q-tilde, E q-tilde) generations of each */
static int g_n = 0;
static long g_gen = -1;                 /* l-space generation the caches belong to */
static long g_egen = -1;                /* E setup generation the caches belong to (-1: none) */
static int g_prim_nblocks = 1;          /* thread count of the last decompose_tetrapyd_prim */

static void drop_caches(void)
{
    free(g_gamma3d); g_gamma3d = NULL;
    free(g_gammagl); g_gammagl = NULL;
    for (int k = 0; k < 2; ++k) { free(g_cross[k]); g_cross[k] = NULL; g_cross_gen[k] = -1; }
    for (int k = 0; k < 4; ++k) { free(g_pol[k]); g_pol[k] = NULL; }
    for (int sp = 0; sp < 2; ++sp) {
        for (int k = 0; k < 4; ++k) { free(g_crossp[sp][k]); g_crossp[sp][k] = NULL; }
        g_crossp_gen[sp][0] = g_crossp_gen[sp][1] = -1;
    }
}

int mcmb_legacy_bind(mcmb_ctx* ctx)
{
    pthread_mutex_lock(&g_mu);
    g_ctx = ctx;
    drop_caches();
    g_n = 0;
    g_gen = -1;
    g_egen = -1;
    g_status = MCMB_OK;
    pthread_mutex_unlock(&g_mu);
    return MCMB_OK;
}

/* Refresh sizes from the bound context; drop caches of an older setup.  Called
 * with g_mu held. */
static int refresh(void)
{
    int lmin, lmax, n;
    long gen;
    if (!g_ctx) return MCMB_ERR_STATE;
    int st = mcmb_lspace_info(g_ctx, &lmin, &lmax, &n, &gen);
    if (st != MCMB_OK) return st;
    long egen = -1;
    if (mcmb_lspace_E_info(g_ctx, NULL, NULL, NULL, &egen) != MCMB_OK) egen = -1;
    if (gen != g_gen || egen != g_egen) {
        drop_caches();
        g_gen = gen;
        g_egen = egen;
    }
    g_lsize = lmax + 1;
    g_n = n;
    return MCMB_OK;
}

/* lsize of the walk of the 3D Gamma: get_qtilde_T_lsize() of the q-tilde
 * setup when there is one, else lmax + 1 of the l-space setup. */
static int walk_lsize(void)
{
    int lq;
    if (g_ctx && mcmb_qtilde_info(g_ctx, NULL, &lq, NULL, NULL, NULL) == MCMB_OK) return lq + 1;
    return g_lsize;
}

int mcmb_legacy_status(void) { return g_status; }

/* ---- tetrapyd walks ------------------------------------------------------ */

/* tetrapyd_tools.c:104-142: next point of {i<=j<=k<=lmax, k<=i+j, even sum};
 * 0 when the step lands on the final point. */
int get_ijk_next_XXX(int lmax, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk;
    int odd = i & 1;
    if (j + odd == lmax) { ++i; j = i; k = j + (i & 1); }
    else if (k + 2 > lmax || k >= i + j) { ++j; k = j + odd; }
    else k += 2;
    int more = 1;
    if (i + j + k + 1 >= 3 * lmax) {
        more = 0;
        i = (lmax & 1) ? lmax - 1 : lmax;
        j = k = lmax;
    }
    *pi = i; *pj = j; *pk = k;
    return more;
}

static void order3(int* a, int* b, int* c)
{
    int t;
    if (*a > *b) { t = *a; *a = *b; *b = t; }
    if (*b > *c) { t = *b; *b = *c; *c = t; if (*a > *b) { t = *a; *a = *b; *b = t; } }
}

/* tetrapyd_tools.c:144-180 */
void check_ijk_min_XXX(int lmin, int* pi, int* pj, int* pk)
{
    int i = *pi < lmin ? lmin : *pi, j = *pj < lmin ? lmin : *pj, k = *pk < lmin ? lmin : *pk;
    order3(&i, &j, &k);
    if (k > i + j) k = i + j;
    else if ((i + j + k) & 1) ++k;
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:182-224 */
void check_ijk_max_XXX(int lmax, int* pi, int* pj, int* pk)
{
    int i = *pi > lmax ? lmax : *pi, j = *pj > lmax ? lmax : *pj, k = *pk > lmax ? lmax : *pk;
    order3(&i, &j, &k);
    if (k > i + j) k = i + j;
    else if ((i + j + k) & 1) {
        if (k > j) --k;
        else if (j > i) --j;
        else --i;
    }
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:595-608 */
void divide_tasks(long int tasks, int nranks, long int* per)
{
    long int q = tasks / nranks, r = tasks % nranks;
    for (int w = 0; w < nranks; ++w) per[w] = q + (w < r);
}

/* tetrapyd_tools.c:11-102 */
void decompose_tetrapyd_XXX(tetrapyd_limits* blk, int rank, int nranks, int nthreads, int lmin, int lmax)
{
    int i0 = lmin, j0 = lmin, k0 = lmin;
    check_ijk_min_XXX(lmin, &i0, &j0, &k0);
    int i = i0, j = j0, k = k0;
    long cap = (long)lmax * lmax * lmax + 1, count = 1;
    while (count < cap && get_ijk_next_XXX(lmax, &i, &j, &k)) ++count;
    ++count;
    long* pr = (long*)malloc(sizeof(long) * (size_t)nranks);
    long* pt = (long*)malloc(sizeof(long) * (size_t)nthreads);
    divide_tasks(count, nranks, pr);
    divide_tasks(pr[rank], nthreads, pt);
    i = i0; j = j0; k = k0;
    for (int r = 0; r < rank; ++r)
        for (long s = 0; s < pr[r]; ++s) get_ijk_next_XXX(lmax, &i, &j, &k);
    for (int t = 0; t < nthreads; ++t) {
        if (t)
            for (long s = 0; s < pt[t - 1]; ++s) get_ijk_next_XXX(lmax, &i, &j, &k);
        blk[t].i_bgn = i; blk[t].j_bgn = j; blk[t].k_bgn = k; blk[t].loops = (int)pt[t];
    }
    free(pr);
    free(pt);
}

/* tetrapyd_tools.c:299-337: the mixed walk, i <= j over the X range and k
 * over the Y range, k from max(j - i, lminY) (parity-bumped) to min(i + j,
 * lmaxY) in steps of 2.  Kept as the reference has it: a new i restarts at
 * k = lminY with no parity bump, a j whose k range is empty still yields a
 * point, and the walk ends on i + j + k + 1 >= 2 lmaxX + lmaxY at the point
 * (lmaxX, lmaxX, lmaxY rounded down to even), which for an odd lmaxY is
 * reached early, from (lmaxX - 1, lmaxX, lmaxY).  It is the mixed tetrapyd
 * exactly when lminX = lminY and lmaxX = lmaxY, both even. */
int get_ijk_next_XXY(int lminX, int lminY, int lmaxX, int lmaxY, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk;
    const int ktop = i + j < lmaxY ? i + j : lmaxY;
    if (j == lmaxX && k + 2 > lmaxY) {
        ++i;
        j = i;
        k = lminY;   /* (i + j) % 2 == 0 */
    } else if (k + 2 > ktop) {
        ++j;
        k = j - i > lminY ? j - i : lminY;
        if ((i + j + k) & 1) ++k;
    } else {
        k += 2;
    }
    if (i + j + k + 1 >= 2 * lmaxX + lmaxY) {
        *pi = lmaxX; *pj = lmaxX; *pk = lmaxY - (lmaxY & 1);
        return 0;
    }
    *pi = i; *pj = j; *pk = k;
    return 1;
}

/* tetrapyd_tools.c:339-381 */
void check_ijk_min_XXY(int lminX, int lminY, int* pi, int* pj, int* pk)
{
    int i = *pi < lminX ? lminX : *pi, j = *pj < lminX ? lminX : *pj, k = *pk < lminY ? lminY : *pk;
    if (i > j) { int t = i; i = j; j = t; }
    if (k > i + j) {
        if (i + j >= lminY) k = i + j;
        else if (lminY > 2 * lminX) { i = lminY / 2; j = i + (lminY & 1); k = lminY; }
        else { i = j = lminX; k = lminY + (lminY & 1); }
    } else if (j > i + k) {
        j = i + k;
    } else if ((i + j + k) & 1) {
        ++k;
    }
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:383-414 */
void check_ijk_max_XXY(int lmaxX, int lmaxY, int* pi, int* pj, int* pk)
{
    int i = *pi > lmaxX ? lmaxX : *pi, j = *pj > lmaxX ? lmaxX : *pj, k = *pk > lmaxY ? lmaxY : *pk;
    if (i > j) { int t = i; i = j; j = t; }
    if (k > i + j) k = i + j;
    else if (j > i + k) j = i + k;
    else if ((i + j + k) & 1) --k;
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:226-297: count the walk (capped at lmaxX^2 lmaxY steps),
 * split it over ranks and then threads, walk forward to each slice start */
void decompose_tetrapyd_XXY(tetrapyd_limits* blk, int rank, int nranks, int nthreads, int lminX, int lminY,
                            int lmaxX, int lmaxY)
{
    int i0 = lminX, j0 = lminX, k0 = lminY;
    check_ijk_min_XXY(lminX, lminY, &i0, &j0, &k0);
    int i = i0, j = j0, k = k0;
    long cap = (long)lmaxX * lmaxX * lmaxY + 1, count = 1;
    while (count < cap && get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k)) ++count;
    ++count;
    long* pr = (long*)malloc(sizeof(long) * (size_t)nranks);
    long* pt = (long*)malloc(sizeof(long) * (size_t)nthreads);
    divide_tasks(count, nranks, pr);
    divide_tasks(pr[rank], nthreads, pt);
    i = i0; j = j0; k = k0;
    for (int r = 0; r < rank; ++r)
        for (long s = 0; s < pr[r]; ++s) get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k);
    for (int t = 0; t < nthreads; ++t) {
        if (t)
            for (long s = 0; s < pt[t - 1]; ++s) get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, &i, &j, &k);
        blk[t].i_bgn = i; blk[t].j_bgn = j; blk[t].k_bgn = k; blk[t].loops = (int)pt[t];
    }
    free(pr);
    free(pt);
}

/* tetrapyd_tools.c:485-517 */
int get_ijk_next_prim(int max, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk;
    if (j == max) { ++i; j = i; k = j; }
    else if (k == max || k >= i + j + 2) { ++j; k = j; }
    else ++k;
    int more = 1;
    if (i >= max) { more = 0; i = j = k = max; }
    *pi = i; *pj = j; *pk = k;
    return more;
}

/* tetrapyd_tools.c:519-554 */
void check_ijk_min_prim(int min, int* pi, int* pj, int* pk)
{
    int i = *pi < min ? min : *pi, j = *pj < min ? min : *pj, k = *pk < min ? min : *pk;
    order3(&i, &j, &k);
    if (k > i + j + 2) k = i + j + 2;
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:556-592 */
void check_ijk_max_prim(int max, int* pi, int* pj, int* pk)
{
    int i = *pi > max ? max : *pi, j = *pj > max ? max : *pj, k = *pk > max ? max : *pk;
    order3(&i, &j, &k);
    if (k > i + j + 2) { k = i + j + 2; if (k > max) k = max; }
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:416-483 */
void decompose_tetrapyd_prim(tetrapyd_limits* blk, int rank, int nranks, int nthreads, int min, int max)
{
    g_prim_nblocks = nthreads > 0 ? nthreads : 1;
    int i0 = min, j0 = min, k0 = min;
    check_ijk_min_prim(min, &i0, &j0, &k0);
    int i = i0, j = j0, k = k0;
    long cap = (long)(max + 1) * (max + 1) * (max + 1), count = 1;
    while (count < cap && get_ijk_next_prim(max, &i, &j, &k)) ++count;
    ++count;
    long* pr = (long*)malloc(sizeof(long) * (size_t)nranks);
    long* pt = (long*)malloc(sizeof(long) * (size_t)nthreads);
    divide_tasks(count, nranks, pr);
    divide_tasks(pr[rank], nthreads, pt);
    i = i0; j = j0; k = k0;
    for (int r = 0; r < rank; ++r)
        for (long s = 0; s < pr[r]; ++s) get_ijk_next_prim(max, &i, &j, &k);
    for (int t = 0; t < nthreads; ++t) {
        if (t)
            for (long s = 0; s < pt[t - 1]; ++s) get_ijk_next_prim(max, &i, &j, &k);
        blk[t].i_bgn = i; blk[t].j_bgn = j; blk[t].k_bgn = k; blk[t].loops = (int)pt[t];
    }
    free(pr);
    free(pt);
}

/* calculate_gamma_3d.c:213-272: advance up to maxblocksize points through
 * T(2, lsize-1) from (i0,j0,k0); (i1,j1,k1) is the exclusive end, capped at
 * the final bound; *blocksize is the number of points passed. */
void mcmb_next_ijk_bound(int ls, int i0, int j0, int k0, int* i1, int* j1, int* k1,
                         int fi, int fj, int fk, int* blocksize, int maxblocksize)
{
    int oi = fi, oj = fj, ok = fk, cnt = 0;
    for (int i = i0; i <= fi; ++i) {
        int jlo = (i == i0) ? j0 : i;
        int jhi = (i == fi) ? fj + 1 : ls;
        for (int j = jlo; j < jhi; ++j) {
            int klo = (i == i0 && j == j0) ? k0 : j + (i % 2);
            int kmx = (i + j < ls - 1) ? i + j : ls - 1;
            int khi = (i == fi && j == fj) ? fk : kmx + 1;
            for (int k = klo; k < khi; k += 2) {
                if (cnt == maxblocksize) { oi = i; oj = j; ok = k; goto out; }
                ++cnt;
            }
        }
    }
out:
    if (oi >= fi && oj >= fj && ok > fk) { oi = fi; oj = fj; ok = fk; }
    *i1 = oi; *j1 = oj; *k1 = ok;
    *blocksize = cnt;
}

void get_next_ijk_bound(int i0, int j0, int k0, int* i1, int* j1, int* k1,
                        int fi, int fj, int fk, int* blocksize, int maxblocksize)
{
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    int ls = walk_lsize();
    pthread_mutex_unlock(&g_mu);
    if (st != MCMB_OK) { *i1 = fi; *j1 = fj; *k1 = fk; *blocksize = 0; g_status = MCMB_ERR_STATE; return; }
    mcmb_next_ijk_bound(ls, i0, j0, k0, i1, j1, k1, fi, fj, fk, blocksize, maxblocksize);
}

/* calculate_gamma_3d.c:92-211, the host build: the walk's point count split
 * evenly over the ranks (the first count % numranks ranks take one more), the
 * rank's host share (1 - workshare) of it, then split over the threads the
 * same way; thread t gets [i0,j0,k0] .. [i1,j1,k1) in td[t].  The thread
 * count is omp_get_num_threads() of the caller's team when the program links
 * OpenMP (a weak reference), else 1. */
extern int omp_get_num_threads(void) __attribute__((weak));

void mcmb_decompose_gamma_TTT_3D(int lsize, int rank, int numranks, double workshare, int nthreads,
                                 thread_g3d_data_t* td)
{
    long cnt = 0;
    for (int i = 2; i < lsize; ++i)
        for (int j = i; j < lsize; ++j) {
            int ke = i + j < lsize - 1 ? i + j : lsize - 1;
            int kb = j + (i % 2);
            if (kb <= ke) cnt += (ke - kb) / 2 + 1;
        }
    int count = (int)cnt;   /* the reference counts in int */
    int blk = count / numranks, left = count - blk * numranks, dummy;
    int i0 = 2, j0 = 2, k0 = 2, i1, j1, k1;
    for (int r = 0; r < rank; ++r) {
        mcmb_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, lsize, lsize, lsize, &dummy, blk + (r < left));
        i0 = i1; j0 = j1; k0 = k1;
    }
    int work = blk + (rank < left);
    int host = (int)((1.0 - workshare) * work);
    mcmb_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, lsize, lsize, lsize, &dummy, host);
    int fi = i1, fj = j1, fk = k1;
    int tb = host / nthreads, tl = host - tb * nthreads;
    for (int t = 0; t < nthreads; ++t) {
        mcmb_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, fi, fj, fk, &dummy, tb + (t < tl));
        td[t].i0 = i0; td[t].j0 = j0; td[t].k0 = k0;
        td[t].i1 = i1; td[t].j1 = j1; td[t].k1 = k1;
        i0 = i1; j0 = j1; k0 = k1;
    }
}

void decompose_gamma_TTT_3D_mpi(thread_g3d_data_t* td, int rank, int numranks, double workshare)
{
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    int ls = walk_lsize();
    pthread_mutex_unlock(&g_mu);
    if (st != MCMB_OK || numranks < 1 || rank < 0 || rank >= numranks) {
        g_status = st != MCMB_OK ? st : MCMB_ERR_ARG;
        return;
    }
    int nt = omp_get_num_threads ? omp_get_num_threads() : 1;
    mcmb_decompose_gamma_TTT_3D(ls, rank, numranks, workshare, nt > 0 ? nt : 1, td);
    g_status = MCMB_OK;
}

/* ---- weights and basis -------------------------------------------------- */

/* fixed_data.c:4244-4254 */
double permsix(int i, int j, int k)
{
    if (i == j && j == k) return 1.0;
    return (i == j || j == k || i == k) ? 3.0 : 6.0;
}

/* calculate_geometric.c:59-92 */
double calculate_geometric(int l1, int l2, int l3)
{
    if (((l1 + l2 + l3) & 1) || l1 + l2 < l3 || l2 + l3 < l1 || l3 + l1 < l2) return 0.0;
    double s1 = l1 + l2 + l3 + 1.0, s3 = l1 + l2 + l3 + 1.0 / 3.0, s6 = l1 + l2 + l3 + 1.0 / 6.0;
    double t13 = l2 + l3 - l1 + 1.0 / 3.0, t23 = l3 + l1 - l2 + 1.0 / 3.0, t33 = l1 + l2 - l3 + 1.0 / 3.0;
    double t16 = l2 + l3 - l1 + 1.0 / 6.0, t26 = l3 + l1 - l2 + 1.0 / 6.0, t36 = l1 + l2 - l3 + 1.0 / 6.0;
    double c = M_1_PI * M_1_PI / 2.0;
    double p1 = c * ((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0)) / s1;
    double p2 = s3 / (t13 * t23 * t33);
    double p3 = sqrt(t16 * t26 * t36 / s6);
    return p1 * p2 * p3;
}

/* basis_functions.c:88-107 */
int basis_functions_legendre(double** a, int size, int pmax, double min, double max, double* values)
{
    for (int i = 0; i < size; ++i) {
        double x = (values[i] - min) / (max - min);
        a[i][0] = 1.0;
        if (pmax > 0) a[i][1] = 2.0 * x - 1.0;
        for (int j = 2; j <= pmax; ++j)
            a[i][j] = ((2.0 * j - 1.0) * a[i][1] * a[i][j - 1] - (j - 1.0) * a[i][j - 2]) / (1.0 * j);
    }
    return 0;
}

/* ---- Gram entry points ---------------------------------------------------- */

/* gamma_mpi.c:30-152: zeroes the caller's buffer and fills it with this
 * rank's share of the 3D cross-Gamma of the bound context (late basis of the
 * l-space setup against q-tilde), 6 dphi^2 included; summing the buffers of
 * all ranks gives the whole, as the MPI_Reduce at gamma_mpi.c:764 does.
 * workshare (the Xeon Phi offload split) is accepted and ignored.  Needs
 * mcmb_qtilde_set and gamma_size = both mode counts (MCMB_ERR_STATE /
 * MCMB_ERR_ARG in mcmb_legacy_status otherwise). */
void gamma_3d_host(double* gamma_flat, int gamma_size, int rank, int numranks, double workshare)
{
    (void)workshare;
    memset(gamma_flat, 0, sizeof(double) * (size_t)gamma_size * gamma_size);
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    pthread_mutex_unlock(&g_mu);
    int nq = 0;
    if (st == MCMB_OK) st = mcmb_qtilde_info(g_ctx, NULL, NULL, NULL, &nq, NULL) == MCMB_OK ? MCMB_OK : MCMB_ERR_STATE;
    if (st != MCMB_OK) { g_status = st; return; }
    if (gamma_size != g_n || gamma_size != nq) { g_status = MCMB_ERR_ARG; return; }
    g_status = mcmb_gamma_cross_3d(g_ctx, rank, numranks > 0 ? numranks : 1, gamma_flat, NULL);
}

/* The whole Gram of the bound context in *slot (kind 0: FP64 tetrapyd, 1: GL),
 * computed on first use; g_mu held. */
static int cached(double** slot, int gl)
{
    int st0 = refresh();
    if (st0 != MCMB_OK) return st0;
    if (*slot) return MCMB_OK;
    double* g = (double*)malloc(sizeof(double) * (size_t)g_n * g_n);
    if (!g) return MCMB_ERR_NOMEM;
    int st = gl ? mcmb_gamma_gl(g_ctx, 0, g, NULL) : mcmb_gamma_tet(g_ctx, MCMB_FP64, g, NULL);
    if (st != MCMB_OK) { free(g); return st; }
    *slot = g;
    return MCMB_OK;
}

/* Entry (m, n) of a cached Gram; status */
static int cached_entry(double** slot, int gl, int m, int n, double* out)
{
    pthread_mutex_lock(&g_mu);
    int st = cached(slot, gl);
    int nn = g_n;
    double v = 0.0;
    if (st == MCMB_OK && (m < 0 || n < 0 || m >= nn || n >= nn)) st = MCMB_ERR_ARG;
    if (st == MCMB_OK) v = (*slot)[(size_t)m * nn + n];
    pthread_mutex_unlock(&g_mu);
    *out = v;
    return st;
}

/* orthogonalisation_late.c:338 (TTT only) */
int calculate_orthol_3D(int m, int n, double* result)
{
    g_status = cached_entry(&g_gamma3d, 0, m, n, &result[0]);
    return g_status;
}

/* orthogonalisation_late.c:239-336: result[0] = Gamma_TTT(m, n) of the GL
 * Gram; with an E setup bound (mcmb_lspace_set_E, the reference's
 * do_polarisation = 1) also result[1..3] = TTE, TEE, EEE, every component on
 * 3 max(T lmax, E lmax)/2 + 1 nodes as init_orthol_glint takes them then.
 * Each component is computed whole on first use and cached. */
int calculate_orthol(int m, int n, double* result)
{
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    const int pol = st == MCMB_OK && g_egen >= 0;
    int ne = 0;
    if (pol && mcmb_lspace_E_info(g_ctx, NULL, NULL, &ne, NULL) == MCMB_OK && ne != g_n) st = MCMB_ERR_ARG;
    if (st == MCMB_OK && (m < 0 || n < 0 || m >= g_n || n >= g_n)) st = MCMB_ERR_ARG;
    if (st == MCMB_OK && !pol) {
        pthread_mutex_unlock(&g_mu);
        g_status = cached_entry(&g_gammagl, 1, m, n, &result[0]);
        return g_status;
    }
    for (int k = 0; k < 4 && st == MCMB_OK; ++k) {
        if (!g_pol[k]) {
            double* g = (double*)malloc(sizeof(double) * (size_t)g_n * g_n);
            if (!g) { st = MCMB_ERR_NOMEM; break; }
            st = mcmb_gamma_gl_pol(g_ctx, k, 0, g, NULL);
            if (st != MCMB_OK) { free(g); break; }
            g_pol[k] = g;
        }
        result[k] = g_pol[k][(size_t)m * g_n + n];
    }
    pthread_mutex_unlock(&g_mu);
    g_status = st;
    return st;
}

/* orthogonalisation_late.c:23-97: the GL tables are built by the first
 * calculate_orthol on the bound context; returns 0 as the reference does. */
double init_orthol_glint(void)
{
    pthread_mutex_lock(&g_mu);
    g_status = refresh();
    pthread_mutex_unlock(&g_mu);
    return 0.0;
}

/* orthogonalisation_late.c:99-129: M_ij of every (i, j) is formed on the GPU
 * in one pass when the Gram is first needed; this call only checks state. */
void calculate_orthoMij_T(int i, int j)
{
    (void)i;
    (void)j;
    pthread_mutex_lock(&g_mu);
    g_status = refresh();
    pthread_mutex_unlock(&g_mu);
}

/* orthogonalisation_late.c:131-157: as calculate_orthoMij_T for the E field;
 * MCMB_ERR_STATE without an E setup */
void calculate_orthoMij_E(int i, int j)
{
    (void)i;
    (void)j;
    pthread_mutex_lock(&g_mu);
    g_status = refresh();
    if (g_status == MCMB_OK && g_egen < 0) g_status = MCMB_ERR_STATE;
    pthread_mutex_unlock(&g_mu);
}

/* orthogonalisation_late.c:159: one GPU holds every M_ij, nothing to sum */
void sync_orthoMij(void) { g_status = MCMB_OK; }

/* ---- GL cross-Gamma (calculate_gamma.c) ------------------------------------ */

/* Entry (m, n) of the cached GL cross-Gamma with x rule spline (0 FMM, 1
 * natural) into result[0]; with the polarised setups bound (mcmb_lspace_set_E
 * and mcmb_qtilde_set_E: the reference's do_polarisation = 1) into
 * result[0..3] = TTT, TTE, TEE, EEE.  Each is computed whole on first use
 * under g_mu and kept until a setup changes. */
static int cross_entry(int spline, int m, int n, double* result)
{
    pthread_mutex_lock(&g_mu);
    int nq = 0, ne = 0;
    long qgen = -1, qegen = -1;
    int st = refresh();
    if (st == MCMB_OK && mcmb_qtilde_info(g_ctx, NULL, NULL, NULL, &nq, &qgen) != MCMB_OK) st = MCMB_ERR_STATE;
    const int pol = st == MCMB_OK && g_egen >= 0 && mcmb_qtilde_E_info(g_ctx, NULL, &qegen) == MCMB_OK;
    if (pol && mcmb_lspace_E_info(g_ctx, NULL, NULL, &ne, NULL) == MCMB_OK && ne != g_n) st = MCMB_ERR_ARG;
    if (st == MCMB_OK && (m < 0 || n < 0 || m >= g_n || n >= nq)) st = MCMB_ERR_ARG;
    if (st == MCMB_OK && !pol) {
        if (g_cross[spline] && g_cross_gen[spline] != qgen) {
            free(g_cross[spline]);
            g_cross[spline] = NULL;
        }
        if (!g_cross[spline]) {
            double* g = (double*)malloc(sizeof(double) * (size_t)g_n * nq);
            if (!g) st = MCMB_ERR_NOMEM;
            else if ((st = mcmb_gamma_cross_gl(g_ctx, spline, 0, g, NULL)) != MCMB_OK) free(g);
            else { g_cross[spline] = g; g_cross_gen[spline] = qgen; }
        }
        if (st == MCMB_OK) result[0] = g_cross[spline][(size_t)m * nq + n];
    } else if (st == MCMB_OK) {
        if (g_crossp_gen[spline][0] != qgen || g_crossp_gen[spline][1] != qegen) {
            for (int k = 0; k < 4; ++k) { free(g_crossp[spline][k]); g_crossp[spline][k] = NULL; }
            g_crossp_gen[spline][0] = qgen;
            g_crossp_gen[spline][1] = qegen;
        }
        for (int k = 0; k < 4 && st == MCMB_OK; ++k) {
            if (!g_crossp[spline][k]) {
                double* g = (double*)malloc(sizeof(double) * (size_t)g_n * nq);
                if (!g) { st = MCMB_ERR_NOMEM; break; }
                if ((st = mcmb_gamma_cross_gl_pol(g_ctx, k, spline, 0, g, NULL)) != MCMB_OK) { free(g); break; }
                g_crossp[spline][k] = g;
            }
            result[k] = g_crossp[spline][k][(size_t)m * nq + n];
        }
    }
    pthread_mutex_unlock(&g_mu);
    return st;
}

/* calculate_gamma.c:772-843: result[0] = Gamma_TTT(m, n), FMM spline in x
 * (result[1..3] = TTE, TEE, EEE with the polarised setups bound) */
void calculate_gamma_2d(int m, int n, double* result)
{
    g_status = cross_entry(0, m, n, result);
}

/* calculate_gamma.c:936-1016: the same with the natural spline in x */
void calculate_gamma(int m, int n, double* result)
{
    g_status = cross_entry(1, m, n, result);
}

/* calculate_gamma.c:48-240: the tables are built on the GPU by the first
 * calculate_gamma / calculate_gamma_2d; this call checks the setups. */
void init_gamma_glint(void)
{
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    if (st == MCMB_OK && mcmb_qtilde_info(g_ctx, NULL, NULL, NULL, NULL, NULL) != MCMB_OK) st = MCMB_ERR_STATE;
    pthread_mutex_unlock(&g_mu);
    g_status = st;
}

/* calculate_gamma.c:255-303: forms every M_ij and the FMM-spline Gamma whole */
void precompute_gammaMij_T(void)
{
    double v[4];
    g_status = cross_entry(0, 0, 0, v);
}

/* calculate_gamma.c:306-352: the E tables are formed with the T ones, by the
 * first cross-Gamma of the polarised setups; MCMB_ERR_STATE without them */
void precompute_gammaMij_E(void)
{
    double v[4];
    pthread_mutex_lock(&g_mu);
    int st = refresh();
    if (st == MCMB_OK && (g_egen < 0 || mcmb_qtilde_E_info(g_ctx, NULL, NULL) != MCMB_OK)) st = MCMB_ERR_STATE;
    pthread_mutex_unlock(&g_mu);
    g_status = st == MCMB_OK ? cross_entry(0, 0, 0, v) : st;
}

/* calculate_gamma.c:354-386: per-(i, j) M of one rank; the GPU forms them all */
void calculate_gammaMij_T(int i, int j)
{
    (void)i;
    (void)j;
    init_gamma_glint();
}

/* calculate_gamma.c:388-423: as calculate_gammaMij_T for the E field */
void calculate_gammaMij_E(int i, int j)
{
    (void)i;
    (void)j;
    init_gamma_glint();
    if (g_status != MCMB_OK) return;
    pthread_mutex_lock(&g_mu);
    if (g_egen < 0 || mcmb_qtilde_E_info(g_ctx, NULL, NULL) != MCMB_OK) g_status = MCMB_ERR_STATE;
    pthread_mutex_unlock(&g_mu);
}

/* calculate_gamma.c:425-509: one GPU holds every M_ij, nothing to sum */
void sync_gammaMij(void) { g_status = MCMB_OK; }

/* fixed_data.c get_qtilde_xsize / get_qtilde_xvec / get_qtilde_T_lsize of the
 * bound q-tilde setup (0 / untouched without one) */
int get_qtilde_xsize(void)
{
    int nx = 0;
    if (!g_ctx || mcmb_qtilde_info(g_ctx, NULL, NULL, &nx, NULL, NULL) != MCMB_OK) return 0;
    return nx;
}

int get_qtilde_xvec(double* xvec)
{
    g_status = g_ctx ? mcmb_qtilde_xvec(g_ctx, xvec) : MCMB_ERR_STATE;
    return g_status;
}

int get_qtilde_T_lsize(void)
{
    int lq = 0;
    if (!g_ctx || mcmb_qtilde_info(g_ctx, NULL, &lq, NULL, NULL, NULL) != MCMB_OK) return 0;
    return lq + 1;
}

/* Index of (i, j, k) in the prim walk from (min, min, min): rows (i, j) in
 * order hold k in [j, min(max, i+j+2)]. */
static long prim_walk_index(int min, int max, int i, int j, int k)
{
    long p = 0;
    for (int a = min; a <= i; ++a)
        for (int b = a; b <= max; ++b) {
            if (a == i && b == j) return p + (k - j);
            int khi = a + b + 2 < max ? a + b + 2 : max;
            p += khi - b + 1;
        }
    return -1;
}

/* Point range of the blocks of the last decompose_tetrapyd_prim: block[0]'s
 * start and the loops of all its threads (one contiguous share of the walk). */
static int prim_block_range(int min, int max, const tetrapyd_limits* block, long* p0, long* p1)
{
    int umin, umax, n;
    double kmax;
    if (!g_ctx) return MCMB_ERR_STATE;
    int st = mcmb_kuniform_info(g_ctx, &umin, &umax, &n, &kmax, NULL);
    if (st != MCMB_OK) return st;
    if (umin != min || umax != max) return MCMB_ERR_ARG;
    *p0 = prim_walk_index(min, max, block[0].i_bgn, block[0].j_bgn, block[0].k_bgn);
    if (*p0 < 0) return MCMB_ERR_ARG;
    long loops = 0;
    for (int t = 0; t < g_prim_nblocks; ++t) loops += block[t].loops;
    *p1 = *p0 + loops;
    return MCMB_OK;
}

/* orthogonalisation.c:14 -- adds the Gram of the blocks' share of the
 * uniform-grid walk into the packed upper triangle ortho[n(n+1)/2] (row
 * major, r <= s), integrating every block (single-thread semantics). */
void calculate_ortho(int min, int max, tetrapyd_limits* block, double* ortho)
{
    long p0, p1;
    int n;
    g_status = prim_block_range(min, max, block, &p0, &p1);
    if (g_status != MCMB_OK) return;
    mcmb_kuniform_info(g_ctx, NULL, NULL, &n, NULL, NULL);
    double* g = (double*)malloc(sizeof(double) * (size_t)n * n);
    if (!g) { g_status = MCMB_ERR_NOMEM; return; }
    g_status = mcmb_kuniform_gram(g_ctx, MCMB_FP64, p0, p1, g, NULL);
    if (g_status == MCMB_OK) {
        long m = 0;
        for (int r = 0; r < n; ++r)
            for (int c = r; c < n; ++c) ortho[m++] += g[(long)r * n + c];
    }
    free(g);
}

/* calculate_eigen.c:86 -- adds alpha of the feature shape (model 4;
 * params.a1 = omega, params.a2 = phase in percent of 2 pi as
 * calculate_shape.c:933-946 reads it) over the blocks' share of the walk. */
void calculate_eigen(int min, int max, shape_params params, tetrapyd_limits* block, double* alpha)
{
    long p0, p1;
    int n;
    g_status = prim_block_range(min, max, block, &p0, &p1);
    if (g_status != MCMB_OK) return;
    mcmb_kuniform_info(g_ctx, NULL, NULL, &n, NULL, NULL);
    double* a = (double*)malloc(sizeof(double) * (size_t)n);
    if (!a) { g_status = MCMB_ERR_NOMEM; return; }
    double theta[2] = {params.a1, 2.0 * 3.1415927 * params.a2 * 1e-2};
    g_status = mcmb_kuniform_project(g_ctx, 1, theta, p0, p1, a);
    if (g_status == MCMB_OK)
        for (int r = 0; r < n; ++r) alpha[r] += a[r];
    free(a);
}

/* ---- reference helpers over the bound context's tables ------------------ */

/* orthogonalisation.c:279-295 -- pijk of the uniform k-grid setup (basis row =
 * grid index, order table of mcmb_kuniform_set); 0 if unset or out of range. */
double pijk(int r, int i, int j, int k)
{
    const double* q;
    const int* order;
    int ldq, n, mx;
    if (!g_ctx || mcmb_kuniform_tables(g_ctx, &q, &ldq, &order, &n, &mx) != MCMB_OK) return 0.0;
    if (r < 0 || r >= n || i < 0 || j < 0 || k < 0 || i > mx || j > mx || k > mx) return 0.0;
    const int p1 = order[3 * r], p2 = order[3 * r + 1], p3 = order[3 * r + 2];
    const double* qi = q + (long)i * ldq;
    const double* qj = q + (long)j * ldq;
    const double* qk = q + (long)k * ldq;
    double b1 = qi[p1] * qj[p2] * qk[p3];
    double b2 = qi[p2] * qj[p3] * qk[p1];
    double b3 = qi[p3] * qj[p1] * qk[p2];
    double b4 = qi[p3] * qj[p2] * qk[p1];
    double b5 = qi[p2] * qj[p1] * qk[p3];
    double b6 = qi[p1] * qj[p3] * qk[p2];
    return (b1 + b2 + b3 + b4 + b5 + b6) / (6.0);
}

/* orthogonalisation_late.c:620-644 -- the same product over the l-space setup
 * (basis row = multipole l), 0 outside [lmin, lmax]. */
double plijk_TTT(int r, int l1, int l2, int l3)
{
    const double* q;
    const int* order;
    int ldq, n, lmin, lmax;
    if (!g_ctx || mcmb_lspace_tables(g_ctx, &q, &ldq, &order, &n, &lmin, &lmax) != MCMB_OK) return 0.0;
    if (r < 0 || r >= n) return 0.0;
    if (l1 < lmin || l1 > lmax || l2 < lmin || l2 > lmax || l3 < lmin || l3 > lmax) return 0.0;
    const int p1 = order[3 * r], p2 = order[3 * r + 1], p3 = order[3 * r + 2];
    const double* qi = q + (long)l1 * ldq;
    const double* qj = q + (long)l2 * ldq;
    const double* qk = q + (long)l3 * ldq;
    double b1 = qi[p1] * qj[p2] * qk[p3];
    double b2 = qi[p2] * qj[p3] * qk[p1];
    double b3 = qi[p3] * qj[p1] * qk[p2];
    double b4 = qi[p3] * qj[p2] * qk[p1];
    double b5 = qi[p2] * qj[p1] * qk[p3];
    double b6 = qi[p1] * qj[p3] * qk[p2];
    return (b1 + b2 + b3 + b4 + b5 + b6) / (6e0);
}

/* fixed_data.c:263-270 */
void find_perm_prim(int r, int* p1, int* p2, int* p3)
{
    const double* q;
    const int* order;
    int ldq, n, mx;
    *p1 = *p2 = *p3 = -1;
    if (!g_ctx || mcmb_kuniform_tables(g_ctx, &q, &ldq, &order, &n, &mx) != MCMB_OK || r < 0 || r >= n) return;
    *p1 = order[3 * r];
    *p2 = order[3 * r + 1];
    *p3 = order[3 * r + 2];
}

/* fixed_data.c:983-990 -- the late order table of the bound l-space setup */
void find_perm_late_TTT(int r, int* p1, int* p2, int* p3)
{
    const double* q;
    const int* order;
    int ldq, n, lmin, lmax;
    *p1 = *p2 = *p3 = -1;
    if (!g_ctx || mcmb_lspace_tables(g_ctx, &q, &ldq, &order, &n, &lmin, &lmax) != MCMB_OK || r < 0 || r >= n) return;
    *p1 = order[3 * r];
    *p2 = order[3 * r + 1];
    *p3 = order[3 * r + 2];
}

/* fixed_data.c:992-1019 -- mode r of the bound E setup's TTE / TEE / EEE
 * table, in the slot order given to mcmb_lspace_set_E (-1s outside it) */
static void find_perm_late_E(int table, int r, int* p1, int* p2, int* p3)
{
    const int* t[3] = {NULL, NULL, NULL};
    int n = 0;
    *p1 = *p2 = *p3 = -1;
    if (!g_ctx || mcmb_lspace_E_tables(g_ctx, NULL, NULL, &t[0], &t[1], &t[2], &n, NULL, NULL) != MCMB_OK || r < 0 ||
        r >= n)
        return;
    *p1 = t[table][3 * r];
    *p2 = t[table][3 * r + 1];
    *p3 = t[table][3 * r + 2];
}
void find_perm_late_TTE(int r, int* p1, int* p2, int* p3) { find_perm_late_E(0, r, p1, p2, p3); }
void find_perm_late_TEE(int r, int* p1, int* p2, int* p3) { find_perm_late_E(1, r, p1, p2, p3); }
void find_perm_late_EEE(int r, int* p1, int* p2, int* p3) { find_perm_late_E(2, r, p1, p2, p3); }

/* orthogonalisation_late.c:646-715 -- the mixed mode products over the T
 * (l-space) and E setups, basis rows by multipole; 0 when a leg is outside
 * its field's range:
 *   plijk_TTE = (qT_p1(l1) qT_p2(l2) + qT_p2(l1) qT_p1(l2)) qE_p3(l3) / 2,
 *   plijk_TEE = qT_p1(l1) (qE_p2(l2) qE_p3(l3) + qE_p3(l2) qE_p2(l3)) / 2,
 *   plijk_EEE = the six-permutation mean over the E basis. */
static double plijk_pol(int table, int r, int l1, int l2, int l3)
{
    const double *qT, *qE;
    const int* t[3] = {NULL, NULL, NULL};
    int ldT, ldE, nT, nE, tlo, thi, elo, ehi;
    if (!g_ctx || mcmb_lspace_tables(g_ctx, &qT, &ldT, NULL, &nT, &tlo, &thi) != MCMB_OK ||
        mcmb_lspace_E_tables(g_ctx, &qE, &ldE, &t[0], &t[1], &t[2], &nE, &elo, &ehi) != MCMB_OK)
        return 0.0;
    if (r < 0 || r >= nE) return 0.0;
    const int* o = t[table] + 3 * r;
    /* field of each leg: TTE (T, T, E), TEE (T, E, E), EEE (E, E, E) */
    const int l[3] = {l1, l2, l3};
    for (int a = 0; a < 3; ++a) {
        const int e = table == 2 || (table == 1 && a > 0) || (table == 0 && a == 2);
        if (e ? (l[a] < elo || l[a] > ehi) : (l[a] < tlo || l[a] > thi)) return 0.0;
    }
    if (table == 0) {
        const double* a = qT + (long)l1 * ldT;
        const double* b = qT + (long)l2 * ldT;
        return (a[o[0]] * b[o[1]] + a[o[1]] * b[o[0]]) * qE[(long)l3 * ldE + o[2]] / 2.0;
    }
    if (table == 1) {
        const double* b = qE + (long)l2 * ldE;
        const double* c = qE + (long)l3 * ldE;
        return qT[(long)l1 * ldT + o[0]] * (b[o[1]] * c[o[2]] + b[o[2]] * c[o[1]]) / 2.0;
    }
    const double* a = qE + (long)l1 * ldE;
    const double* b = qE + (long)l2 * ldE;
    const double* c = qE + (long)l3 * ldE;
    const int p1 = o[0], p2 = o[1], p3 = o[2];
    return (a[p1] * b[p2] * c[p3] + a[p2] * b[p3] * c[p1] + a[p3] * b[p1] * c[p2] + a[p3] * b[p2] * c[p1] +
            a[p2] * b[p1] * c[p3] + a[p1] * b[p3] * c[p2]) / 6.0;
}
double plijk_TTE(int r, int l1, int l2, int l3) { return plijk_pol(0, r, l1, l2, l3); }
double plijk_TEE(int r, int l1, int l2, int l3) { return plijk_pol(1, r, l1, l2, l3); }
double plijk_EEE(int r, int l1, int l2, int l3) { return plijk_pol(2, r, l1, l2, l3); }

/* fixed_data.c get_terms_prim / get_kmax of the uniform setup (0 if unset) */
int get_terms_prim(void)
{
    int n = 0;
    if (!g_ctx || mcmb_kuniform_info(g_ctx, NULL, NULL, &n, NULL, NULL) != MCMB_OK) return 0;
    return n;
}

double get_kmax(void)
{
    double kmax = 0.0;
    if (!g_ctx || mcmb_kuniform_info(g_ctx, NULL, NULL, NULL, &kmax, NULL) != MCMB_OK) return 0.0;
    return kmax;
}

/* calculate_shape.c:933-946 (model 4, the feature shape): sin(a1 K + 2 pi a2 %)
 * / (k1 k2 k3)^2, 0 when k1 k2 k3 <= 1e-15; pi as the reference writes it. */
double shape(double k1, double k2, double k3, shape_params params)
{
    double p1 = k1 + k2 + k3, p111 = k1 * k2 * k3;
    double factor = 0.0;
    if (p111 > 1e-15) factor = 1.0 / (p111 * p111);
    return factor * sin(params.a1 * p1 + 2.0 * 3.1415927 * params.a2 * 1e-2);
}

/* calculate_shape.c:197-216 */
double shape3(double k1, double k2, double k3, shape_params params)
{
    double p222 = (k1 * k1) * (k2 * k2) * (k3 * k3);
    return p222 * shape(k1, k2, k3, params);
}

/* fixed_data.c:1670-1715 for the Legendre orders (order_prim < 4): rebuild the
 * uniform k-grid setup of the bound context on size points of [0, max] with
 * basis q[i][p] = P_p(2 vec[i]/max - 1) (basis_functions_legendre with min 0)
 * for p up to the largest index of its order table.  Returns 0 (the
 * reference's value) or a negative status. */
int create_basis_prim(int size, double max, double* vec)
{
    const double* q0;
    const int* order;
    int ldq, n, mx;
    if (!g_ctx) return MCMB_ERR_STATE;
    int st = mcmb_kuniform_tables(g_ctx, &q0, &ldq, &order, &n, &mx);
    if (st != MCMB_OK) return st;
    if (size < 2) return MCMB_ERR_ARG;
    int pmax = 0;
    for (int e = 0; e < 3 * n; ++e) if (order[e] > pmax) pmax = order[e];
    int* ord = (int*)malloc(sizeof(int) * 3 * (size_t)n);
    double* q = (double*)malloc(sizeof(double) * (size_t)size * (pmax + 1));
    double** rows = (double**)malloc(sizeof(double*) * (size_t)size);
    if (!ord || !q || !rows) { free(ord); free(q); free(rows); return MCMB_ERR_NOMEM; }
    memcpy(ord, order, sizeof(int) * 3 * (size_t)n);
    for (int i = 0; i < size; ++i) rows[i] = q + (long)i * (pmax + 1);
    basis_functions_legendre(rows, size, pmax, 0.0, max, vec);
    st = mcmb_kuniform_set(g_ctx, 0, size - 1, max, pmax + 1, q, n, ord);
    free(rows);
    free(q);
    free(ord);
    return st;
}
