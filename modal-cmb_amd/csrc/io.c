/* File layouts of the reference's hot-path outputs and order tables.
 *
 * Gamma, ortho, lambda and eigen files (output_gamma, output_ortho,
 * output_lambda, output_eigen: fixed_data.c:3077-3091, 3297-3310, 3588-3601,
 * 2899-2928) are written by array_write / vector_write (arrays.c:6-21) as raw
 * native-endian FP64 with no header: n x n row-major matrices, length-n
 * vectors.  Order tables (load_three_int, io.c:229-258) are text: one triple
 * per line, each field parsed as a double and truncated to int, lines
 * starting with '#' skipped.  Unlike the reference, these functions return a
 * status instead of exiting, and a short read is an error. */
#include "modal_cmb.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MCMB_MAXLINES 10000000   /* reference MAXLINES (global.h:7) */
#define MCMB_MAXLEN 200          /* reference MAXLEN (global.h:6) */

int mcmb_write_raw(const char* path, long count, const double* data)
{
    if (!path || count < 0 || (count > 0 && !data)) return MCMB_ERR_ARG;
    FILE* fp = fopen(path, "wb");
    if (!fp) return MCMB_ERR_IO;
    size_t w = count ? fwrite(data, sizeof(double), (size_t)count, fp) : 0;
    int bad = fclose(fp) != 0 || w != (size_t)count;
    return bad ? MCMB_ERR_IO : MCMB_OK;
}

int mcmb_read_raw(const char* path, long count, double* data)
{
    if (!path || count < 0 || (count > 0 && !data)) return MCMB_ERR_ARG;
    FILE* fp = fopen(path, "rb");
    if (!fp) return MCMB_ERR_IO;
    size_t r = count ? fread(data, sizeof(double), (size_t)count, fp) : 0;
    fclose(fp);
    return r == (size_t)count ? MCMB_OK : MCMB_ERR_IO;
}

int mcmb_load_order(const char* path, int* order, int cap, int* nmodes)
{
    if (!path || !order || cap < 1 || !nmodes) return MCMB_ERR_ARG;
    FILE* fp = fopen(path, "r");
    if (!fp) return MCMB_ERR_IO;
    char line[MCMB_MAXLEN];
    int n = 0, st = MCMB_OK;
    while (fgets(line, sizeof(line), fp)) {
        if (*line == '#') continue;
        char* p = line;
        char* q;
        double d[3];
        int ok = 1;
        for (int a = 0; a < 3; ++a) {
            d[a] = strtod(p, &q);
            if (q == p) { ok = 0; break; }
            p = q;
        }
        if (!ok) {
            /* a blank line is skipped; anything else is malformed */
            for (char* c = line; *c; ++c)
                if (*c != ' ' && *c != '\t' && *c != '\n' && *c != '\r') { st = MCMB_ERR_ORDER; break; }
            if (st != MCMB_OK) break;
            continue;
        }
        if (n >= cap) { st = MCMB_ERR_ARG; break; }
        order[3 * n] = (int)d[0];
        order[3 * n + 1] = (int)d[1];
        order[3 * n + 2] = (int)d[2];
        ++n;
    }
    fclose(fp);
    *nmodes = n;
    return st;
}

/* io.c:229-258 -- returns 1 on success, 0 if the file cannot be read; at
 * most MAXLINES triples are stored (values must hold 3 * that many). */
int load_three_int(char* filename, int* values, int* size)
{
    int n = 0;
    int st = mcmb_load_order(filename, values, MCMB_MAXLINES, &n);
    *size = n;
    return st == MCMB_OK ? 1 : 0;
}
