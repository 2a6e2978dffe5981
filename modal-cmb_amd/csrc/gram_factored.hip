// Factored tetrapyd Gram: the same Gamma_mn = sum_p w_p Q_m(p) Q_n(p) over the
// rows of the ordered tetrapyd, reorganised through the separable basis.
//
// Within a row (fixed i, j; k = k0 + kstep t) the mode product is
//   Q_m(i,j,k) = (1/6) [ P[rs] q_t(k) + P[rt] q_s(k) + P[st] q_r(k) ],
//   P[xy] = q_x(i) q_y(j) + q_y(i) q_x(j)          ((r,s,t) = order[m]),
// so with the row moments  M_row[cd] = sum_t w q_c(k_t) q_d(k_t)  (c <= d)
//   Gamma_mn = (1/36) sum_{a,b} Z[{p_a, p_b}][{c_a, d_b}],
//   Z[p1][p2][cd] = sum_rows P_row[p1] P_row[p2] M_row[cd],
// where (p_a, c_a) runs over the three (pair, single) splits of mode m and
// (p_b, d_b) over those of mode n.  Z is symmetric in p1 <-> p2 and in c <-> d
// separately, so only p1 <= p2 and c <= d are formed.  With NP = R(R+1)/2
// pairs this is one GEMM of (NP x NP) pair-pair rows by NP columns over
// K = rows: 2 NP^3/2 x rows flops instead of 2 n^2/2 x points for the direct
// contraction (3.8e14 instead of 1.6e16 at lmax 2500, 4960 modes).
//
// Kernels: row moments (FP64, one wave per row); the Z product on FP64 MFMA
// (v_mfma_f64_16x16x4f64), tiles of 8 p1 x 16 p2 rows by 128 cd columns, the
// A operand P[p1] P[p2] formed from LDS-staged row coefficients when it is
// loaded, M staged through LDS double-buffered; K split into contiguous row
// shares summed in a fixed order; assembly of Gamma from Z (9 terms each).
#include "mcmb_internal.h"
#include <algorithm>
#include <cstdlib>
#include <vector>

namespace mcmb {

typedef double fd4 __attribute__((ext_vector_type(4)));

constexpr int FZ_P1 = 8;      // p1 values per tile
constexpr int FZ_P2 = 16;     // p2 values per tile
constexpr int FZ_TN = 128;    // cd columns per tile
constexpr int FZ_KC = 16;     // rows per K chunk
constexpr int FZ_MAXU = 9;    // pairs per lane in the row-moment kernel (NP <= 576)

// M[row][cd] = sum_t w_t q_c(k_t) q_d(k_t), one wave per row, lanes over pairs.
__global__ __launch_bounds__(256) void fz_rowmom(const RowDesc* __restrict__ rows, const long* __restrict__ rowptr,
                                                 int row_lo, int nrows, const double* __restrict__ q, int ldq,
                                                 int kstep, const double* __restrict__ ws,
                                                 const int2* __restrict__ pxy, int NP, int NPp,
                                                 double* __restrict__ M)
{
    const int lr = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (lr >= nrows) return;
    const int lane = threadIdx.x & 63;
    const RowDesc rd = rows[row_lo + lr];
    const long base = rowptr[row_lo + lr];
    int xs[FZ_MAXU], ys[FZ_MAXU];
    double acc[FZ_MAXU];
#pragma unroll
    for (int u = 0; u < FZ_MAXU; ++u) {
        const int pp = lane + 64 * u;
        const int2 xy = pp < NP ? pxy[pp] : make_int2(0, 0);
        xs[u] = xy.x;
        ys[u] = xy.y;
        acc[u] = 0.0;
    }
    for (int t = 0; t < rd.nk; ++t) {
        const double* qk = q + (long)(rd.k0 + kstep * t) * ldq;
        const double w = ws[base + t];
#pragma unroll
        for (int u = 0; u < FZ_MAXU; ++u) acc[u] += w * qk[xs[u]] * qk[ys[u]];
    }
#pragma unroll
    for (int u = 0; u < FZ_MAXU; ++u) {
        const int pp = lane + 64 * u;
        if (pp < NP) M[(long)lr * NPp + pp] = acc[u];
    }
}

// Zpart[split][p1][p2][cd] = sum over this split's rows of P[p1] P[p2] M[cd].
// Workgroup: 4 waves, 2 x 2 grid of 64 x 64 (A rows = (p1, p2), columns cd).
__global__ __launch_bounds__(256, 2) void fz_gemm_f64(
    const RowDesc* __restrict__ rows, int row_lo, int nrows, const double* __restrict__ q, int ldq,
    const double* __restrict__ M, int NPp, const int2* __restrict__ pxy, int NP,
    const int2* __restrict__ tiles, int ncd, int nsplit, int rows_per_split, double* __restrict__ Zpart)
{
    __shared__ double Ms[2][FZ_KC][FZ_TN + 2];
    __shared__ double P1s[2][FZ_KC][FZ_P1];
    __shared__ double P2s[2][FZ_KC][FZ_P2];
    const int cdb = blockIdx.x % ncd;
    const int split = (blockIdx.x / ncd) % nsplit;
    const int tile = blockIdx.x / (ncd * nsplit);
    const int2 ab = tiles[tile];
    const int a0 = ab.x, b0 = ab.y, cd0 = cdb * FZ_TN;
    const int r_lo = split * rows_per_split;
    const int r_hi = min(r_lo + rows_per_split, nrows);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    fd4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (fd4){0.0, 0.0, 0.0, 0.0};

    double rm[8], rp[2];
    auto gload = [&](int r0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + 256 * u;                 // 16 rows x 128 columns
            const int kk = e >> 7, cc = e & 127;
            const int r = r0 + kk, cd = cd0 + cc;
            rm[u] = (r < r_hi && cd < NP) ? M[(long)r * NPp + cd] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {                    // 16 rows x (8 + 16) pair values
            const int e = tid + 256 * u;
            double v = 0.0;
            if (e < FZ_KC * (FZ_P1 + FZ_P2)) {
                const int kk = e / (FZ_P1 + FZ_P2), s = e % (FZ_P1 + FZ_P2);
                const int p = s < FZ_P1 ? a0 + s : b0 + (s - FZ_P1);
                const int r = r0 + kk;
                if (r < r_hi && p < NP) {
                    const RowDesc rd = rows[row_lo + r];
                    const int2 xy = pxy[p];
                    const double* qi = q + (long)rd.i * ldq;
                    const double* qj = q + (long)rd.j * ldq;
                    v = qi[xy.x] * qj[xy.y] + qi[xy.y] * qj[xy.x];
                }
            }
            rp[u] = v;
        }
    };
    auto sstore = [&](int b) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + 256 * u;
            Ms[b][e >> 7][e & 127] = rm[u];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + 256 * u;
            if (e < FZ_KC * (FZ_P1 + FZ_P2)) {
                const int kk = e / (FZ_P1 + FZ_P2), s = e % (FZ_P1 + FZ_P2);
                if (s < FZ_P1) P1s[b][kk][s] = rp[u];
                else P2s[b][kk][s - FZ_P1] = rp[u];
            }
        }
    };

    const int nch = (r_hi - r_lo + FZ_KC - 1) / FZ_KC;
    if (nch > 0) {
        gload(r_lo);
        sstore(0);
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
            if (ch + 1 < nch) gload(r_lo + (ch + 1) * FZ_KC);
#pragma unroll
            for (int ks = 0; ks < FZ_KC / 4; ++ks) {
                const int kk = ks * 4 + (lane >> 4);
                const double p2 = P2s[b][kk][lane & 15];
                double av[4], bv[4];
#pragma unroll
                for (int a = 0; a < 4; ++a) av[a] = P1s[b][kk][(wm * 4 + a) & 7] * p2;
#pragma unroll
                for (int c = 0; c < 4; ++c) bv[c] = Ms[b][kk][wn * 64 + c * 16 + (lane & 15)];
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int c = 0; c < 4; ++c)
                        acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
            }
            if (ch + 1 < nch) sstore(b ^ 1);
            __syncthreads();
        }
    }
    // A rows wm*64 + a*16 + (4v + lane/16): p1 = a0 + wm*4 + a, p2 = b0 + 4v + lane/16
    double* const Zs = Zpart + (size_t)split * NP * NP * NP;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        const int p1 = a0 + wm * 4 + a;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int cd = cd0 + wn * 64 + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int p2 = b0 + 4 * v + (lane >> 4);
                if (p1 < NP && p2 < NP && cd < NP) Zs[((size_t)p1 * NP + p2) * NP + cd] = acc[a][c][v];
            }
        }
    }
}

// Z = sum over splits (in split order) of Zpart, into split 0.
__global__ void fz_reduce(double* __restrict__ Zpart, long nz, int nsplit)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nz) return;
    double s = Zpart[e];
    for (int k = 1; k < nsplit; ++k) s += Zpart[(long)k * nz + e];
    Zpart[e] = s;
}

// Gamma[m][n] = (1/36) sum over the 3 x 3 (pair, single) splits of Z entries.
__global__ void fz_assemble(const double* __restrict__ Z, int NP, int R, const int* __restrict__ modes, int n,
                            double* __restrict__ gamma)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int m = (int)(e / n), nn = (int)(e % n);
    if (m > nn) return;
    int pa[3], ca[3], pb[3], cb[3];
    {
        const int pk = modes[m];
        const int r = pk & 255, s = (pk >> 8) & 255, t = (pk >> 16) & 255;
        const int u[3][3] = {{r, s, t}, {r, t, s}, {s, t, r}};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            pa[a] = pair_index(min(u[a][0], u[a][1]), max(u[a][0], u[a][1]), R);
            ca[a] = u[a][2];
        }
    }
    {
        const int pk = modes[nn];
        const int r = pk & 255, s = (pk >> 8) & 255, t = (pk >> 16) & 255;
        const int u[3][3] = {{r, s, t}, {r, t, s}, {s, t, r}};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            pb[a] = pair_index(min(u[a][0], u[a][1]), max(u[a][0], u[a][1]), R);
            cb[a] = u[a][2];
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const int p1 = min(pa[a], pb[b]), p2 = max(pa[a], pb[b]);
            const int cd = pair_index(min(ca[a], cb[b]), max(ca[a], cb[b]), R);
            acc += Z[((size_t)p1 * NP + p2) * NP + cd];
        }
    acc *= (1.0 / 36.0);
    gamma[(long)m * n + nn] = acc;
    gamma[(long)nn * n + m] = acc;
}

bool gram_factored_supported(const PointSet& L, int precision)
{
    return precision == MCMB_FP64 && L.R * (L.R + 1) / 2 <= 64 * FZ_MAXU;
}

int gram_factored_launch(mcmb_ctx* c, PointSet& L, int precision, double* d_gamma, int row_lo, int row_hi)
{
    if (precision != MCMB_FP64) return fail(MCMB_ERR_ARG, "factored Gram: precision %d not supported", precision);
    const int n = L.n, R = L.R;
    const int NP = R * (R + 1) / 2, NPp = NP;
    if (NP > 64 * FZ_MAXU) return fail(MCMB_ERR_ARG, "factored Gram: basis width %d too large", R);
    const int nrows = row_hi - row_lo;
    if (nrows <= 0) {
        MCMB_HIP(hipMemsetAsync(d_gamma, 0, sizeof(double) * (size_t)n * n, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        c->last_ms = 0.0;
        c->last_launches = 0;
        return MCMB_OK;
    }
    std::vector<int2> pxy(NP);
    for (int x = 0; x < R; ++x)
        for (int y = x; y < R; ++y) pxy[pair_index(x, y, R)] = make_int2(x, y);
    std::vector<int2> tiles;
    for (int a0 = 0; a0 < NP; a0 += FZ_P1)
        for (int b0 = 0; b0 < NP; b0 += FZ_P2)
            if (a0 <= b0 + FZ_P2 - 1) tiles.push_back(make_int2(a0, b0));
    const int ncd = (NP + FZ_TN - 1) / FZ_TN;
    const long nz = (long)NP * NP * NP;
    // split the rows so that the grid fills the chip (>= 1024 workgroups),
    // with the partial Z buffers within 4 GiB
    int nsplit = (int)std::max<long>(1, (1024 + (long)tiles.size() * ncd - 1) / ((long)tiles.size() * ncd));
    nsplit = (int)std::min<long>(nsplit, std::max<long>(1, (4l << 30) / (nz * 8)));
    nsplit = std::min(nsplit, std::max(1, nrows / FZ_KC));
    const int rows_per_split = (nrows + nsplit - 1) / nsplit;

    DevBuf rowptr, ws, d_pxy, d_tiles, d_M;
    MCMB_TRY(pointset_weights(c, L, rowptr, ws, 0));
    MCMB_TRY(d_pxy.alloc(sizeof(int2) * pxy.size()));
    MCMB_TRY(d_tiles.alloc(sizeof(int2) * tiles.size()));
    MCMB_TRY(d_M.alloc(sizeof(double) * (size_t)nrows * NPp));
    MCMB_TRY(c->d_partial.alloc(sizeof(double) * (size_t)nsplit * nz));
    MCMB_HIP(hipMemcpyAsync(d_pxy.p, pxy.data(), sizeof(int2) * pxy.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    MCMB_TRY(launch_guard("row moments", (unsigned long long)((nrows + 3) / 4), 256));
    hipLaunchKernelGGL(fz_rowmom, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, c->stream, L.d_rows.as<RowDesc>(),
                       rowptr.as<long>(), row_lo, nrows, L.d_q.as<double>(), L.ldq, L.kstep, ws.as<double>(),
                       d_pxy.as<int2>(), NP, NPp, d_M.as<double>());
    MCMB_HIP(hipGetLastError());
    const unsigned long long nwg = (unsigned long long)tiles.size() * ncd * nsplit;
    MCMB_TRY(launch_guard("factored Gram product", nwg, 256));
    hipLaunchKernelGGL(fz_gemm_f64, dim3((unsigned)nwg), dim3(256), 0, c->stream, L.d_rows.as<RowDesc>(), row_lo,
                       nrows, L.d_q.as<double>(), L.ldq, d_M.as<double>(), NPp, d_pxy.as<int2>(), NP,
                       d_tiles.as<int2>(), ncd, nsplit, rows_per_split, c->d_partial.as<double>());
    MCMB_HIP(hipGetLastError());
    if (nsplit > 1) {
        hipLaunchKernelGGL(fz_reduce, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, c->stream,
                           c->d_partial.as<double>(), nz, nsplit);
        MCMB_HIP(hipGetLastError());
    }
    const long ng = (long)n * n;
    hipLaunchKernelGGL(fz_assemble, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, c->stream,
                       c->d_partial.as<double>(), NP, R, L.d_modes.as<int>(), n, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 3 + (nsplit > 1);
    return MCMB_OK;
}

}  // namespace mcmb
