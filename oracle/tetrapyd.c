/* Tetrapyd walks, decompositions and block bounds -- CPU restatement (test
 * infrastructure only, see oracle.h). */
#include "oracle.h"
#include <stdlib.h>

/* Step to the next point of the even-parity, triangle-limited ordered set
 * {i<=j<=k<=lmax, k<=i+j, i+j+k even}: tetrapyd_tools.c:104-142.
 * The i row ends when j reaches lmax - i%2 (the only k left there is lmax);
 * a j row ends when the next k would pass lmax or k already equals i+j.
 * Stepping into the final point (sum >= 3 lmax - 1) returns 0 and leaves the
 * triple on that final point, which depends on the parity of lmax. */
int or_get_ijk_next_XXX(int lmax, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk, ok = 1;
    if (j + (i & 1) == lmax) {
        i += 1;
        j = i;
        k = j + (i & 1);
    } else if (k >= i + j || k + 2 > lmax) {
        j += 1;
        k = j + (i & 1);
    } else {
        k += 2;
    }
    if (i + j + k + 1 >= 3 * lmax) {
        ok = 0;
        i = (lmax % 2 == 0) ? lmax : lmax - 1;
        j = lmax;
        k = lmax;
    }
    *pi = i; *pj = j; *pk = k;
    return ok;
}

static void sort3(int* a, int* b, int* c)
{
    int t;
    if (*a > *b) { t = *a; *a = *b; *b = t; }
    if (*b > *c) {
        t = *b; *b = *c; *c = t;
        if (*a > *b) { t = *a; *a = *b; *b = t; }
    }
}

/* tetrapyd_tools.c:144-180: clamp up to lmin, order, then move k onto the
 * triangle edge or bump it to even parity. */
void or_check_ijk_min_XXX(int lmin, int* pi, int* pj, int* pk)
{
    int i = *pi < lmin ? lmin : *pi;
    int j = *pj < lmin ? lmin : *pj;
    int k = *pk < lmin ? lmin : *pk;
    sort3(&i, &j, &k);
    if (i + j < k) k = i + j;
    else if ((i + j + k) % 2 == 1) k += 1;
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:182-224: clamp down to lmax, order, then move back onto
 * the triangle edge or step the largest non-tied index down for parity. */
void or_check_ijk_max_XXX(int lmax, int* pi, int* pj, int* pk)
{
    int i = *pi > lmax ? lmax : *pi;
    int j = *pj > lmax ? lmax : *pj;
    int k = *pk > lmax ? lmax : *pk;
    sort3(&i, &j, &k);
    if (i + j < k) {
        k = i + j;
    } else if ((i + j + k) % 2 == 1) {
        if (k > j) k -= 1;
        else if (j > i) j -= 1;
        else i -= 1;
    }
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:595-608 (the first tasks%nranks workers get one extra) */
void or_divide_tasks(long tasks, int nranks, long* per)
{
    long size = tasks / nranks, extra = tasks % nranks;
    for (int r = 0; r < nranks; ++r) per[r] = size + (r < extra ? 1 : 0);
}

/* Point count as tetrapyd_tools.c:31-51 counts it: start at 1, one per
 * successful step, bounded by lmax^3 steps, plus one for the final point. */
long or_count_tetrapyd_XXX(int lmin, int lmax)
{
    int i = lmin, j = lmin, k = lmin;
    or_check_ijk_min_XXX(lmin, &i, &j, &k);
    long limit = (long)lmax * lmax * lmax + 1, count;
    for (count = 1; count < limit; ++count)
        if (!or_get_ijk_next_XXX(lmax, &i, &j, &k)) break;
    return count + 1;
}

long or_walk_XXX(int lmin, int lmax, int* ijk, long cap)
{
    int i = lmin, j = lmin, k = lmin;
    or_check_ijk_min_XXX(lmin, &i, &j, &k);
    long n = or_count_tetrapyd_XXX(lmin, lmax);
    for (long p = 0; p < n; ++p) {
        if (p < cap) { ijk[3 * p] = i; ijk[3 * p + 1] = j; ijk[3 * p + 2] = k; }
        if (p + 1 < n) or_get_ijk_next_XXX(lmax, &i, &j, &k);
    }
    return n;
}

/* tetrapyd_tools.c:11-102: split the counted walk over ranks then threads,
 * and walk forward to each slice start. */
void or_decompose_tetrapyd_XXX(or_tetrapyd_limits* block, int rank, int nranks,
                               int nthreads, int lmin, int lmax)
{
    int i0 = lmin, j0 = lmin, k0 = lmin;
    or_check_ijk_min_XXX(lmin, &i0, &j0, &k0);
    long count = or_count_tetrapyd_XXX(lmin, lmax);
    long* per_rank = (long*)malloc(sizeof(long) * nranks);
    long* per_thread = (long*)malloc(sizeof(long) * nthreads);
    or_divide_tasks(count, nranks, per_rank);
    or_divide_tasks(per_rank[rank], nthreads, per_thread);
    int i = i0, j = j0, k = k0;
    for (int r = 0; r < rank; ++r)
        for (long t = 0; t < per_rank[r]; ++t) or_get_ijk_next_XXX(lmax, &i, &j, &k);
    for (int n = 0; n < nthreads; ++n) {
        if (n > 0)
            for (long t = 0; t < per_thread[n - 1]; ++t) or_get_ijk_next_XXX(lmax, &i, &j, &k);
        block[n].i_bgn = i; block[n].j_bgn = j; block[n].k_bgn = k;
        block[n].loops = (int)per_thread[n];
    }
    free(per_rank);
    free(per_thread);
}

/* tetrapyd_tools.c:485-517: the k-space walk without parity, k up to
 * min(max, i+j+2); the walk ends (returns 0) on reaching i == max. */
int or_get_ijk_next_prim(int max, int* pi, int* pj, int* pk)
{
    int i = *pi, j = *pj, k = *pk, ok = 1;
    if (j == max) {
        i += 1; j = i; k = j;
    } else if (k == max || k >= i + j + 2) {
        j += 1; k = j;
    } else {
        k += 1;
    }
    if (i >= max) { ok = 0; i = max; j = max; k = max; }
    *pi = i; *pj = j; *pk = k;
    return ok;
}

/* tetrapyd_tools.c:519-554 */
void or_check_ijk_min_prim(int min, int* pi, int* pj, int* pk)
{
    int i = *pi < min ? min : *pi;
    int j = *pj < min ? min : *pj;
    int k = *pk < min ? min : *pk;
    sort3(&i, &j, &k);
    if (i + j + 2 < k) k = i + j + 2;
    *pi = i; *pj = j; *pk = k;
}

/* tetrapyd_tools.c:556-592 */
void or_check_ijk_max_prim(int max, int* pi, int* pj, int* pk)
{
    int i = *pi > max ? max : *pi;
    int j = *pj > max ? max : *pj;
    int k = *pk > max ? max : *pk;
    sort3(&i, &j, &k);
    if (i + j + 2 < k) {
        k = i + j + 2;
        if (k > max) k = max;
    }
    *pi = i; *pj = j; *pk = k;
}

static long count_prim(int min, int max)
{
    int i = min, j = min, k = min;
    or_check_ijk_min_prim(min, &i, &j, &k);
    long limit = (long)(max + 1) * (max + 1) * (max + 1), count;
    for (count = 1; count < limit; ++count)
        if (!or_get_ijk_next_prim(max, &i, &j, &k)) break;
    return count + 1;
}

/* tetrapyd_tools.c:416-483 */
void or_decompose_tetrapyd_prim(or_tetrapyd_limits* block, int rank, int nranks,
                                int nthreads, int min, int max)
{
    int i0 = min, j0 = min, k0 = min;
    or_check_ijk_min_prim(min, &i0, &j0, &k0);
    long count = count_prim(min, max);
    long* per_rank = (long*)malloc(sizeof(long) * nranks);
    long* per_thread = (long*)malloc(sizeof(long) * nthreads);
    or_divide_tasks(count, nranks, per_rank);
    or_divide_tasks(per_rank[rank], nthreads, per_thread);
    int i = i0, j = j0, k = k0;
    for (int r = 0; r < rank; ++r)
        for (long t = 0; t < per_rank[r]; ++t) or_get_ijk_next_prim(max, &i, &j, &k);
    for (int n = 0; n < nthreads; ++n) {
        if (n > 0)
            for (long t = 0; t < per_thread[n - 1]; ++t) or_get_ijk_next_prim(max, &i, &j, &k);
        block[n].i_bgn = i; block[n].j_bgn = j; block[n].k_bgn = k;
        block[n].loops = (int)per_thread[n];
    }
    free(per_rank);
    free(per_thread);
}

long or_walk_prim(int min, int max, int* ijk, long cap)
{
    int i = min, j = min, k = min;
    or_check_ijk_min_prim(min, &i, &j, &k);
    long n = count_prim(min, max);
    for (long p = 0; p < n; ++p) {
        if (p < cap) { ijk[3 * p] = i; ijk[3 * p + 1] = j; ijk[3 * p + 2] = k; }
        if (p + 1 < n) or_get_ijk_next_prim(max, &i, &j, &k);
    }
    return n;
}

/* calculate_gamma_3d.c:213-272: walk up to maxblocksize points of T(2, lsize-1)
 * in nested-loop order from (i0,j0,k0); (i1,j1,k1) is the first point NOT in
 * the block (exclusive end), capped at the final bound. */
void or_get_next_ijk_bound(int lsize, int i0, int j0, int k0, int* i1, int* j1, int* k1,
                           int fi, int fj, int fk, int* blocksize, int maxblocksize)
{
    int io = fi, jo = fj, ko = fk, nb = 0;
    for (int i = i0; i <= fi; ++i) {
        int jb = (i == i0) ? j0 : i;
        int je = (i == fi) ? fj + 1 : lsize;
        for (int j = jb; j < je; ++j) {
            int kb = j + (i % 2);
            int ke = (i + j < lsize - 1 ? i + j : lsize - 1) + 1;
            if (i == i0 && j == j0) kb = k0;
            if (i == fi && j == fj) ke = fk;
            for (int k = kb; k < ke; k += 2) {
                if (nb == maxblocksize) { io = i; jo = j; ko = k; goto done; }
                ++nb;
            }
        }
    }
done:
    if (io >= fi && jo >= fj && ko > fk) { io = fi; jo = fj; ko = fk; }
    *i1 = io; *j1 = jo; *k1 = ko;
    *blocksize = nb;
}

/* calculate_gamma_3d.c:92-211, host build (ntasks = nthreads, one worker per
 * task): even split over ranks, host share (1-workshare), then over threads. */
void or_decompose_gamma_TTT_3D(int lsize, int rank, int numranks, double workshare,
                               int nthreads, int* b6)
{
    long count = 0;
    for (int i = 2; i < lsize; ++i)
        for (int j = i; j < lsize; ++j) {
            int ke = (i + j < lsize - 1 ? i + j : lsize - 1);
            for (int k = j + (i % 2); k <= ke; k += 2) ++count;
        }
    int cnt = (int)count;      /* the reference counts in int */
    int blk = cnt / numranks, left = cnt - blk * numranks, dummy;
    int i0 = 2, j0 = 2, k0 = 2, i1, j1, k1;
    for (int r = 0; r < rank; ++r) {
        int w = blk + (r < left ? 1 : 0);
        or_get_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, lsize, lsize, lsize, &dummy, w);
        i0 = i1; j0 = j1; k0 = k1;
    }
    int work = blk + (rank < left ? 1 : 0);
    int host_work = (int)((1.0 - workshare) * work);
    or_get_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, lsize, lsize, lsize, &dummy, host_work);
    work = host_work;
    int rblk = work / nthreads, rleft = work - rblk * nthreads;
    int fi = i1, fj = j1, fk = k1;
    for (int t = 0; t < nthreads; ++t) {
        int w = rblk + (t < rleft ? 1 : 0);
        or_get_next_ijk_bound(lsize, i0, j0, k0, &i1, &j1, &k1, fi, fj, fk, &dummy, w);
        b6[6 * t + 0] = i0; b6[6 * t + 1] = j0; b6[6 * t + 2] = k0;
        b6[6 * t + 3] = i1; b6[6 * t + 4] = j1; b6[6 * t + 5] = k1;
        i0 = i1; j0 = j1; k0 = k1;
    }
}

/* Points p0 .. p0 + cap - 1 of the same walk (fewer at its end); returns the
 * number written.  Walks from the start, as get_ijk_next_XXX must. */
long or_walk_XXX_window(int lmin, int lmax, long p0, int* ijk, long cap)
{
    int i = lmin, j = lmin, k = lmin;
    or_check_ijk_min_XXX(lmin, &i, &j, &k);
    long n = or_count_tetrapyd_XXX(lmin, lmax), w = 0;
    for (long p = 0; p < n && w < cap; ++p) {
        if (p >= p0) { ijk[3 * w] = i; ijk[3 * w + 1] = j; ijk[3 * w + 2] = k; ++w; }
        if (p + 1 < n) or_get_ijk_next_XXX(lmax, &i, &j, &k);
    }
    return w;
}
