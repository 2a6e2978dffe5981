/* The reference's file names and the q-tilde files, host only.
 *
 * File names: the pattern the drivers derive from the .ini (parameter.c:
 * 244-257 for the suffixes, 317-760 for the names): a base name that starts
 * with '/' is used as given (gamma and proj files take a _<pol> / _T, _E
 * suffix); otherwise directory + base + a suffix chain of multipole ranges,
 * basis orders and alpha_max, so that files written here are found by the
 * reference's readers (projecteigen_mpi, fisher_mpi) and the other way round.
 *
 * q-tilde: output_qtilde / read_qtilde (fixed_data.c:2076-2330): a size file
 * of two native ints {lsize + 1, xsize + 1} and, per primordial index p, a
 * data file <data_file>_<p> holding one (lsize + 1) x (xsize + 1) FP64
 * array: [0][0] = p, [0][1 + x] = x grid, [1 + l][0] = multipole of row l,
 * [1 + l][1 + x] = qt[p][l][x].  For primordial order 4 the last two indices
 * are in files _l1 and _l2 instead of _<pmax-1>, _<pmax>. */
#include "modal_cmb.h"
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char* const POL[4] = {"TTT", "TTE", "TEE", "EEE"};

/* append fmt to out (cap bytes); -1 when it does not fit */
static int app(char* out, size_t cap, size_t* len, const char* fmt, ...)
    __attribute__((format(printf, 4, 5)));
static int app(char* out, size_t cap, size_t* len, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int w = vsnprintf(out + *len, cap - *len, fmt, ap);
    va_end(ap);
    if (w < 0 || (size_t)w >= cap - *len) return -1;
    *len += (size_t)w;
    return 0;
}

int mcmb_file_name(char* out, size_t cap, int kind, int pol, const char* directory, const char* base,
                   int T_lmin, int T_lmax, int E_lmin, int E_lmax, int order_prim, int order_late, int alpha_max)
{
    if (!out || cap < 1 || !base || (!directory && base[0] != '/')) return MCMB_ERR_ARG;
    if (pol < 0 || pol > 3) return MCMB_ERR_ARG;
    size_t len = 0;
    out[0] = '\0';
    int bad = 0;
    const int late = kind == MCMB_FILE_ORTHOL || kind == MCMB_FILE_LAMBDAL || kind == MCMB_FILE_GAMMA;
    if (base[0] == '/') {
        bad |= app(out, cap, &len, "%s", base);
        if (kind == MCMB_FILE_GAMMA) bad |= app(out, cap, &len, "_%s", POL[pol]);
        else if (kind == MCMB_FILE_PROJ_DATA) bad |= app(out, cap, &len, pol ? "_E" : "_T");
        else if (kind == MCMB_FILE_PROJ_SIZE) bad |= app(out, cap, &len, pol ? "_E_size" : "_T_size");
        else if (late && pol != 0) return MCMB_ERR_ARG;   /* the reference names only the TTT file */
        return bad ? MCMB_ERR_ARG : MCMB_OK;
    }
    bad |= app(out, cap, &len, "%s%s", directory, base);
    switch (kind) {
    case MCMB_FILE_EIGEN:
    case MCMB_FILE_ORTHO:
    case MCMB_FILE_LAMBDA:
        bad |= app(out, cap, &len, "_%d_%d_%d", T_lmax, order_prim, alpha_max);
        break;
    case MCMB_FILE_ORTHOL:
    case MCMB_FILE_LAMBDAL:
    case MCMB_FILE_GAMMA:
        bad |= app(out, cap, &len, "_%s_", POL[pol]);
        if (pol == 0) bad |= app(out, cap, &len, "%d-%d", T_lmin, T_lmax);
        else if (pol == 3) bad |= app(out, cap, &len, "%d-%d", E_lmin, E_lmax);
        else bad |= app(out, cap, &len, "%d-%d_%d-%d", T_lmin, T_lmax, E_lmin, E_lmax);
        if (kind == MCMB_FILE_GAMMA) bad |= app(out, cap, &len, "_%d", order_prim);
        bad |= app(out, cap, &len, "_%d_%d", order_late, alpha_max);
        break;
    case MCMB_FILE_PROJ_DATA:
        bad |= app(out, cap, &len, pol ? "_E_%d_%d" : "_T_%d_%d", pol ? E_lmax : T_lmax, order_prim);
        break;
    case MCMB_FILE_PROJ_SIZE:
        bad |= app(out, cap, &len, pol ? "_E_%d_size" : "_T_%d_size", pol ? E_lmax : T_lmax);
        break;
    default:
        return MCMB_ERR_ARG;
    }
    return bad ? MCMB_ERR_ARG : MCMB_OK;
}

/* name of the data file of primordial index p of np (order 4: _l1, _l2 for
 * the last two) */
static int qt_file(char* out, size_t cap, const char* data_file, int p, int np, int order4)
{
    int w;
    if (order4 && p == np - 2) w = snprintf(out, cap, "%s_l1", data_file);
    else if (order4 && p == np - 1) w = snprintf(out, cap, "%s_l2", data_file);
    else w = snprintf(out, cap, "%s_%d", data_file, p);
    return (w < 0 || (size_t)w >= cap) ? MCMB_ERR_ARG : MCMB_OK;
}

int mcmb_write_qtilde(const char* size_file, const char* data_file, int np, int order4, int lsize, int xsize,
                      const int* lvec, const double* xvec, const double* qt)
{
    if (!size_file || !data_file || np < 1 || lsize < 1 || xsize < 1 || !lvec || !xvec || !qt) return MCMB_ERR_ARG;
    if (order4 && np < 3) return MCMB_ERR_ARG;
    FILE* fp = fopen(size_file, "wb");
    if (!fp) return MCMB_ERR_IO;
    int sizes[2] = {lsize + 1, xsize + 1};
    size_t wr = fwrite(sizes, sizeof(int), 2, fp);
    if (fclose(fp) != 0 || wr != 2) return MCMB_ERR_IO;
    const long cols = xsize + 1, cnt = (long)(lsize + 1) * cols;
    double* a = (double*)malloc(sizeof(double) * (size_t)cnt);
    if (!a) return MCMB_ERR_NOMEM;
    int st = MCMB_OK;
    char name[4096];
    for (int p = 0; p < np && st == MCMB_OK; ++p) {
        a[0] = (double)p;
        for (int x = 0; x < xsize; ++x) a[1 + x] = xvec[x];
        for (int l = 0; l < lsize; ++l) {
            a[(l + 1) * cols] = (double)lvec[l];
            for (int x = 0; x < xsize; ++x) a[(l + 1) * cols + 1 + x] = qt[((long)p * lsize + l) * xsize + x];
        }
        st = qt_file(name, sizeof(name), data_file, p, np, order4);
        if (st == MCMB_OK) st = mcmb_write_raw(name, cnt, a);
    }
    free(a);
    return st;
}

int mcmb_read_qtilde(const char* size_file, const char* data_file, int np, int order4, int cap_l, int cap_x,
                     int* lsize, int* xsize, int* lvec, double* xvec, double* qt)
{
    if (!size_file || !data_file || np < 1 || !lsize || !xsize) return MCMB_ERR_ARG;
    if (order4 && np < 3) return MCMB_ERR_ARG;
    FILE* fp = fopen(size_file, "rb");
    if (!fp) return MCMB_ERR_IO;
    int sizes[2] = {0, 0};
    size_t rd = fread(sizes, sizeof(int), 2, fp);
    fclose(fp);
    if (rd != 2 || sizes[0] < 2 || sizes[1] < 2) return MCMB_ERR_IO;
    *lsize = sizes[0] - 1;
    *xsize = sizes[1] - 1;
    if (!lvec || !xvec || !qt) return MCMB_OK;   /* sizes only */
    if (*lsize > cap_l || *xsize > cap_x) return MCMB_ERR_ARG;
    const int L = *lsize, X = *xsize;
    const long cols = X + 1, cnt = (long)(L + 1) * cols;
    double* a = (double*)malloc(sizeof(double) * (size_t)cnt);
    if (!a) return MCMB_ERR_NOMEM;
    int st = MCMB_OK;
    char name[4096];
    for (int p = 0; p < np && st == MCMB_OK; ++p) {
        st = qt_file(name, sizeof(name), data_file, p, np, order4);
        if (st == MCMB_OK) st = mcmb_read_raw(name, cnt, a);
        if (st != MCMB_OK) break;
        if (p == 0) {
            for (int x = 0; x < X; ++x) xvec[x] = a[1 + x];
            for (int l = 0; l < L; ++l) lvec[l] = (int)a[(l + 1) * cols];
        }
        for (int l = 0; l < L; ++l)
            for (int x = 0; x < X; ++x) qt[((long)p * L + l) * X + x] = a[(l + 1) * cols + 1 + x];
    }
    free(a);
    return st;
}

/* create_basis_late with eflag_order_late 9 (fixed_data.c:1923-1950), T part:
 * columns 0 .. pmax-2 are the shifted Legendre P_j((l - Tmin)/(Tmax - Tmin))
 * (basis_functions_legendre), column pmax-1 is w/z and column pmax is w z with
 * x = l(l+1)/(Tmax(Tmax+1)), y = (2l+1)/(2 Tmax+1), z = sqrt(x), w = y^(1/6),
 * for the Tsize values Tvec (row i = Tvec[i]).  basis is [Tsize][pmax+1]. */
int mcmb_basis_late_order9(int Tsize, int pmax, double Tmin, double Tmax, const double* Tvec, double* basis)
{
    if (Tsize < 1 || pmax < 2 || !Tvec || !basis || !(Tmax > Tmin)) return MCMB_ERR_ARG;
    const int ld = pmax + 1;
    double** rows = (double**)malloc(sizeof(double*) * (size_t)Tsize);
    if (!rows) return MCMB_ERR_NOMEM;
    for (int i = 0; i < Tsize; ++i) rows[i] = basis + (long)i * ld;
    basis_functions_legendre(rows, Tsize, pmax - 2, Tmin, Tmax, (double*)Tvec);
    free(rows);
    for (int i = 0; i < Tsize; ++i) {
        double* b = basis + (long)i * ld;
        const double x = Tvec[i] * (Tvec[i] + 1.0) / (Tmax * (Tmax + 1.0));
        const double y = (2.0 * Tvec[i] + 1.0) / (2.0 * Tmax + 1.0);
        const double z = sqrt(x), w = pow(y, 1.0 / 6.0);
        b[pmax - 1] = w / z;
        b[pmax] = w * z;
    }
    return MCMB_OK;
}
