// Tetrapyd Gram contraction  Gamma = sum_p w_p Q(p) Q(p)^T  on gfx950.
//
// Work decomposition.  The ordered tetrapyd is a list of rows (i, j) with
// k = k0 + kstep*t, t < nk (walk order of get_ijk_next_XXX).  A row is cut
// into chunks of KC consecutive points; chunks never cross rows, so within a
// chunk the mode product factorises as
//   Q_m(i,j,k) = (1/6) [S_rs q_t(k) + S_rt q_s(k) + S_st q_r(k)],
//   S_ab = q_a(i) q_b(j) + q_b(i) q_a(j)      ((r,s,t) = order[m]),
// three FMAs per (point, mode).  The output is tiled into TM x TM mode tiles
// (I <= J, upper triangle of tiles); the point rows are cut into segments of
// equal chunk count.  One workgroup owns one (tile, segment): it forms the
// operand tiles A[kk][m] = sqrt(w) Q_m / ... in LDS (double buffered, chunk
// c+1 formed while chunk c is contracted) and accumulates A^T B with MFMA in
// registers; per-segment partial tiles are summed in a fixed order by a
// reduction kernel, so results are deterministic.
#include "mcmb_internal.h"
#include "weights.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace mcmb {

constexpr int TM = 128;        // mode tile (FP64 path)
constexpr int KC = 16;         // points per chunk
constexpr int NT = 256;        // threads per workgroup (4 waves)
constexpr int LDA = TM + 4;    // LDS row stride (doubles), staggers banks

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void kspace_weights_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr, long nrows,
    const double* __restrict__ kn, const double* __restrict__ kw, int scaled, double* __restrict__ out)
{
    long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    RowDesc rd = rows[row];
    long base = rowptr[row];
    for (int t = threadIdx.x & 63; t < rd.nk; t += 64) {
        double w = kspace_weight(rd.i, rd.j, rd.k0 + t, kn, kw);
        out[base + t] = scaled ? sqrt(w) * (1.0 / 6.0) : w;
    }
}

// One wave per row: w_p (mode 0) or sqrt(w_p)/6 (mode 1), p in walk order.
__global__ __launch_bounds__(256) void lspace_weights_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr, long nrows, int asym,
    const double* __restrict__ f, const double* __restrict__ v, int scaled,
    double* __restrict__ out)
{
    long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    RowDesc rd = rows[row];
    long base = rowptr[row];
    for (int t = threadIdx.x & 63; t < rd.nk; t += 64) {
        int k = rd.k0 + 2 * t;
        double w = lspace_weight(rd.i, rd.j, k, asym, f, v);
        out[base + t] = scaled ? sqrt(w) * (1.0 / 6.0) : w;
    }
}

// Points [p0, p0 + cap) of the walk: one wave per row, out[p - p0].
__global__ __launch_bounds__(256) void lspace_points_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr, long nrows, long p0, long cap, int kstep,
    int* __restrict__ ijk)
{
    long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    RowDesc rd = rows[row];
    long base = rowptr[row];
    if (base + rd.nk <= p0 || base >= p0 + cap) return;
    for (int t = threadIdx.x & 63; t < rd.nk; t += 64) {
        long p = base + t - p0;
        if (p >= 0 && p < cap) {
            ijk[3 * p] = rd.i;
            ijk[3 * p + 1] = rd.j;
            ijk[3 * p + 2] = rd.k0 + kstep * t;
        }
    }
}

struct ModeInfo {
    int r, s, t;
    bool ok;
};

__device__ __forceinline__ ModeInfo load_mode(const int* __restrict__ modes, int m, int n)
{
    ModeInfo mi;
    if (m < n) {
        int pk = modes[m];
        mi.r = pk & 255; mi.s = (pk >> 8) & 255; mi.t = (pk >> 16) & 255; mi.ok = true;
    } else {
        mi.r = mi.s = mi.t = 0; mi.ok = false;
    }
    return mi;
}

constexpr int QW = 32;   // LDS basis width (ldq <= 32)

// Per-chunk basis staging (LDS): q_c(k) for the KC points [c][kk], q(i), q(j)
// and sqrt(w)/6.  Loaded to registers one chunk ahead (before the MFMA work of
// the current chunk) and stored to LDS after it; branch-free.
struct F64Stage {
    double qk[QW][KC];
    double qi[QW], qj[QW], w[KC];
};
struct F64Regs {
    double q[2];
    double x;
};

__device__ __forceinline__ F64Regs f64_stage_load(const RowDesc& rd, int c0, long pbase, int kstep,
                                                  const double* __restrict__ q, int ldq,
                                                  const double* __restrict__ ws, int tid)
{
    F64Regs g;
    const int last = rd.nk - 1;
#pragma unroll
    for (int u = 0; u < 2; ++u) {             // 512 = QW x KC entries, 2 per thread
        const int e = tid + u * NT, c = e / KC, kk = e % KC, t = c0 + kk;
        const int tc = t < last ? t : last, cc = c < ldq ? c : 0;
        const double v = q[(long)(rd.k0 + kstep * tc) * ldq + cc];
        g.q[u] = (c < ldq && t <= last) ? v : 0.0;
    }
    // tid < 32: q(i)[tid]; 32..63: q(j)[tid-32]; 64..79: w of point c0 + tid-64
    const int c = tid & 31, cc = c < ldq ? c : 0;
    const double qv = q[(long)((tid < 32) ? rd.i : rd.j) * ldq + cc];
    const int kk = tid - 64, t = c0 + (kk < 0 ? 0 : kk);
    const double wv = ws[pbase + ((t <= last) ? (kk < 0 ? 0 : kk) : last - c0)];
    g.x = (tid < 64) ? (c < ldq ? qv : 0.0) : ((kk < KC && t <= last) ? wv : 0.0);
    return g;
}

__device__ __forceinline__ void f64_stage_store(F64Stage& S, const F64Regs& g, int tid)
{
#pragma unroll
    for (int u = 0; u < 2; ++u) (&S.qk[0][0])[tid + u * NT] = g.q[u];
    if (tid < 32) S.qi[tid] = g.x;
    else if (tid < 64) S.qj[tid - 32] = g.x;
    else if (tid < 64 + KC) S.w[tid - 64] = g.x;
}

// Form KC/2 operand values of one mode from the staged chunk into column mloc.
__device__ __forceinline__ void form_f64(const F64Stage& S, const ModeInfo& mi, double* __restrict__ buf,
                                         int mloc, int kb)
{
    double Srs = 0.0, Srt = 0.0, Sst = 0.0;
    if (mi.ok) {
        const double ir = S.qi[mi.r], is = S.qi[mi.s], it = S.qi[mi.t];
        const double jr = S.qj[mi.r], js = S.qj[mi.s], jt = S.qj[mi.t];
        Srs = ir * js + is * jr;
        Srt = ir * jt + it * jr;
        Sst = is * jt + it * js;
    }
#pragma unroll
    for (int e = 0; e < KC / 2; ++e) {
        const int kk = kb + e;
        buf[kk * LDA + mloc] = S.w[kk] * (Srs * S.qk[mi.t][kk] + Srt * S.qk[mi.s][kk] + Sst * S.qk[mi.r][kk]);
    }
}

__global__ __launch_bounds__(NT, 2) void gram_tet_f64_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr,
    const int* __restrict__ seg_rows, int nseg, const int2* __restrict__ tiles, int ntiles,
    const int* __restrict__ modes, int n, const double* __restrict__ q, int ldq, int kstep,
    const double* __restrict__ ws, double* __restrict__ partial, int n_pad)
{
    __shared__ double sA[2][KC * LDA];
    __shared__ double sB[2][KC * LDA];
    __shared__ F64Stage stg[2];
    const int tile = blockIdx.x % ntiles;
    const int seg = blockIdx.x / ntiles;
    const int2 IJ = tiles[tile];
    const bool diag = IJ.x == IJ.y;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int mloc = tid & (TM - 1), kb = (tid >> 7) * (KC / 2);
    const ModeInfo mA = load_mode(modes, IJ.x * TM + mloc, n);
    const ModeInfo mB = load_mode(modes, IJ.y * TM + mloc, n);
    const int wm = wave >> 1, wn = wave & 1;

    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    const int row_begin = seg_rows[seg], row_end = seg_rows[seg + 1];
    long nch = 0;
    for (int r = row_begin; r < row_end; ++r) nch += (rows[r].nk + KC - 1) / KC;
    int prow = row_begin, pc0 = 0;   // next chunk to stage
    auto stage_next = [&](F64Regs& g) {
        const RowDesc rd = rows[prow];
        g = f64_stage_load(rd, pc0, rowptr[prow] + pc0, kstep, q, ldq, ws, tid);
        pc0 += KC;
        if (pc0 >= rd.nk) { ++prow; pc0 = 0; }
    };
    if (nch > 0) {
        F64Regs g;
        stage_next(g);
        f64_stage_store(stg[0], g, tid);
        __syncthreads();
        form_f64(stg[0], mA, sA[0], mloc, kb);
        if (!diag) form_f64(stg[0], mB, sB[0], mloc, kb);
        if (nch > 1) {
            stage_next(g);
            f64_stage_store(stg[1], g, tid);
        }
        __syncthreads();
        for (long ch = 0; ch < nch; ++ch) {
            const int b = (int)(ch & 1);
            const bool more2 = ch + 2 < nch;
            if (more2) stage_next(g);
            const double* A = sA[b];
            const double* B = diag ? sA[b] : sB[b];
#pragma unroll
            for (int ks = 0; ks < KC / 4; ++ks) {
                const int kk = ks * 4 + (lane >> 4);
                double av[4], bv[4];
#pragma unroll
                for (int br = 0; br < 4; ++br) av[br] = A[kk * LDA + wm * 64 + br * 16 + (lane & 15)];
#pragma unroll
                for (int bc = 0; bc < 4; ++bc) bv[bc] = B[kk * LDA + wn * 64 + bc * 16 + (lane & 15)];
#pragma unroll
                for (int br = 0; br < 4; ++br)
#pragma unroll
                    for (int bc = 0; bc < 4; ++bc)
                        acc[br][bc] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[br], bv[bc], acc[br][bc], 0, 0, 0);
            }
            if (ch + 1 < nch) {
                form_f64(stg[b ^ 1], mA, sA[b ^ 1], mloc, kb);
                if (!diag) form_f64(stg[b ^ 1], mB, sB[b ^ 1], mloc, kb);
            }
            if (more2) f64_stage_store(stg[b], g, tid);
            __syncthreads();
        }
    }
    // D layout of v_mfma_f64_16x16x4: lane holds rows 4v + lane/16, column lane%16.
    double* P = partial + (size_t)seg * n_pad * n_pad;
#pragma unroll
    for (int br = 0; br < 4; ++br)
#pragma unroll
        for (int bc = 0; bc < 4; ++bc)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int m = IJ.x * TM + wm * 64 + br * 16 + 4 * v + (lane >> 4);
                int c = IJ.y * TM + wn * 64 + bc * 16 + (lane & 15);
                P[(size_t)m * n_pad + c] = acc[br][bc][v];
            }
}

// Gamma[m][c] = Gamma[c][m] = sum_seg partial[seg][m][c] for m <= c < n.
__global__ __launch_bounds__(256) void gram_reduce_kernel(const double* __restrict__ partial,
                                                          int nseg, int n_pad, int n,
                                                          double* __restrict__ out)
{
    int c = blockIdx.x * 64 + (threadIdx.x & 63);
    int m = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (m >= n || c >= n || m > c) return;
    double s = 0.0;
    size_t off = (size_t)m * n_pad + c, stride = (size_t)n_pad * n_pad;
    for (int g = 0; g < nseg; ++g) s += partial[off + g * stride];
    out[(size_t)m * n + c] = s;
    out[(size_t)c * n + m] = s;
}

// Split rows [lo, hi) into nseg segments of about equal chunk count.
static void make_segments(const std::vector<RowDesc>& rows, int lo, int hi, int kc, int nseg, std::vector<int>& seg)
{
    std::vector<long> pre(hi - lo + 1, 0);
    for (int r = lo; r < hi; ++r) pre[r - lo + 1] = pre[r - lo] + (rows[r].nk + kc - 1) / kc;
    long tot = pre.back();
    seg.assign(nseg + 1, lo);
    for (int s = 1; s < nseg; ++s) {
        long target = tot * s / nseg;
        seg[s] = lo + (int)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
        if (seg[s] < seg[s - 1]) seg[s] = seg[s - 1];
    }
    seg[nseg] = hi;
}

void lspace_slice(const PointSet& L, int kc, int rank, int nranks, int* row_lo, int* row_hi, double* frac)
{
    const std::vector<RowDesc>& rows = L.rows;
    std::vector<long> pre(rows.size() + 1, 0);
    for (size_t r = 0; r < rows.size(); ++r) pre[r + 1] = pre[r] + (rows[r].nk + kc - 1) / kc;
    long tot = pre.back();
    auto at = [&](long target) {
        return (int)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
    };
    *row_lo = rank == 0 ? 0 : at(tot * rank / nranks);
    *row_hi = rank == nranks - 1 ? (int)rows.size() : at(tot * (rank + 1) / nranks);
    if (frac) *frac = tot ? (double)(pre[*row_hi] - pre[*row_lo]) / (double)tot : 0.0;
}

// uniform k-grid weights (kuniform.hip)
int uniform_weights_launch(mcmb_ctx* c, PointSet& L, const long* rowptr, double* ws, int scaled);

static int ensure_lspace_aux(mcmb_ctx* c, PointSet& L, DevBuf& rowptr, DevBuf& ws, int scaled,
                             bool rowptr_ready = false)
{
    if (!rowptr_ready) {
        std::vector<long> rp(L.rows.size() + 1, 0);
        for (size_t r = 0; r < L.rows.size(); ++r) rp[r + 1] = rp[r] + L.rows[r].nk;
        MCMB_TRY(rowptr.alloc(sizeof(long) * rp.size()));
        // synchronous: rp is a local
        MCMB_HIP(hipMemcpy(rowptr.p, rp.data(), sizeof(long) * rp.size(), hipMemcpyHostToDevice));
    }
    MCMB_TRY(ws.alloc(sizeof(double) * (size_t)L.npoints));
    long nrows = (long)L.rows.size();
    dim3 grid((unsigned)((nrows + 3) / 4));
    if (L.uniform)
        return uniform_weights_launch(c, L, rowptr.as<long>(), ws.as<double>(), scaled);
    if (L.kspace)
        hipLaunchKernelGGL(kspace_weights_kernel, grid, dim3(256), 0, c->stream, L.d_rows.as<RowDesc>(),
                           rowptr.as<long>(), nrows, L.d_kn.as<double>(), L.d_kw.as<double>(), scaled,
                           ws.as<double>());
    else
        hipLaunchKernelGGL(lspace_weights_kernel, grid, dim3(256), 0, c->stream, L.d_rows.as<RowDesc>(),
                           rowptr.as<long>(), nrows, L.h2mode == MCMB_H2_ASYMPTOTIC ? 1 : 0,
                           L.d_ftab.as<double>(), L.d_v.as<double>(), scaled, ws.as<double>());
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

int gram_mixed_launch(mcmb_ctx* c, PointSet& L, const long* d_rowptr, const double* d_ws, double* d_gamma,
                      int row_lo, int row_hi);

int pointset_weights(mcmb_ctx* c, PointSet& P, DevBuf& rowptr, DevBuf& ws, int scaled, bool rowptr_ready)
{
    return ensure_lspace_aux(c, P, rowptr, ws, scaled, rowptr_ready);
}

int gram_tet_launch(mcmb_ctx* c, PointSet& L, int precision, double* d_gamma, int row_lo, int row_hi)
{
    c->nphases = 0;
    c->last_rank = -1;
    // FP64: the separable contraction wherever the basis width allows it,
    // else the factored one, else the direct point-by-point contraction below.
    // MCMB_GRAM_PATH = separable | factored | direct selects one explicitly.
    const char* path = getenv("MCMB_GRAM_PATH");
    const bool want_sep = !path || !*path || !strcmp(path, "separable");
    const bool want_fac = !path || !*path || !strcmp(path, "factored");
    if (want_sep && gram_separable_supported(L, precision))
        return gram_separable_launch(c, L, precision, d_gamma, row_lo, row_hi);
    if (want_fac && gram_factored_supported(L, precision))
        return gram_factored_launch(c, L, precision, d_gamma, row_lo, row_hi);
    DevBuf rowptr, ws;
    MCMB_TRY(ensure_lspace_aux(c, L, rowptr, ws, 1));
    if (precision == MCMB_MIXED)
        return gram_mixed_launch(c, L, rowptr.as<long>(), ws.as<double>(), d_gamma, row_lo, row_hi);

    const int n = L.n;
    const int T = (n + TM - 1) / TM, n_pad = T * TM;
    std::vector<int2> tiles;
    for (int I = 0; I < T; ++I)
        for (int J = I; J < T; ++J) tiles.push_back(make_int2(I, J));
    const int ntiles = (int)tiles.size();
    // ~8 workgroups per CU in flight over the launch, bounded by rows and by
    // 4 GiB of partial tiles
    long target = 2048;
    int nseg = (int)std::max<long>(1, (target + ntiles - 1) / ntiles);
    nseg = (int)std::max<long>(1, std::min<long>(nseg, (long)(row_hi - row_lo)));
    size_t per = (size_t)n_pad * n_pad * sizeof(double);
    nseg = (int)std::max<size_t>(1, std::min<size_t>((size_t)nseg, ((size_t)4 << 30) / per));
    std::vector<int> seg;
    make_segments(L.rows, row_lo, row_hi, KC, nseg, seg);

    DevBuf d_tiles, d_seg;
    MCMB_TRY(d_tiles.alloc(sizeof(int2) * tiles.size()));
    MCMB_TRY(d_seg.alloc(sizeof(int) * seg.size()));
    MCMB_TRY(c->d_partial.alloc(per * nseg));
    MCMB_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), sizeof(int2) * tiles.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_seg.p, seg.data(), sizeof(int) * seg.size(), hipMemcpyHostToDevice, c->stream));

    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(gram_tet_f64_kernel, dim3((unsigned)(ntiles * nseg)), dim3(NT), 0, c->stream,
                       L.d_rows.as<RowDesc>(), rowptr.as<long>(), d_seg.as<int>(), nseg,
                       d_tiles.as<int2>(), ntiles, L.d_modes.as<int>(), n, L.d_q.as<double>(), L.ldq, L.kstep,
                       ws.as<double>(), c->d_partial.as<double>(), n_pad);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    dim3 rgrid((unsigned)((n + 63) / 64), (unsigned)((n + 3) / 4));
    hipLaunchKernelGGL(gram_reduce_kernel, rgrid, dim3(256), 0, c->stream, c->d_partial.as<double>(),
                       nseg, n_pad, n, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 1;
    return MCMB_OK;
}

}  // namespace mcmb

using namespace mcmb;

extern "C" long mcmb_lspace_points(mcmb_ctx* c, int* ijk, long cap)
{
    return mcmb_lspace_points_at(c, 0, ijk, cap);
}

extern "C" long mcmb_lspace_points_at(mcmb_ctx* c, long p0, int* ijk, long cap)
{
    if (!c || !c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_points: no l-space setup");
    if (cap < 0 || (cap > 0 && !ijk) || p0 < 0) return fail(MCMB_ERR_ARG, "mcmb_lspace_points: bad buffer or start");
    MCMB_TRY(set_device(c));
    LSpace& L = c->ls;
    long n = L.npoints;
    long m = std::min(cap, std::max(0l, n - p0));
    if (m == 0) return n;
    std::vector<long> rp(L.rows.size() + 1, 0);
    for (size_t r = 0; r < L.rows.size(); ++r) rp[r + 1] = rp[r] + L.rows[r].nk;
    DevBuf rowptr, out;
    MCMB_TRY(rowptr.alloc(sizeof(long) * rp.size()));
    MCMB_TRY(out.alloc(sizeof(int) * 3 * (size_t)m));
    MCMB_HIP(hipMemcpyAsync(rowptr.p, rp.data(), sizeof(long) * rp.size(), hipMemcpyHostToDevice, c->stream));
    long nrows = (long)L.rows.size();
    hipLaunchKernelGGL(lspace_points_kernel, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, c->stream,
                       L.d_rows.as<RowDesc>(), rowptr.as<long>(), nrows, p0, m, L.kstep, out.as<int>());
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipMemcpyAsync(ijk, out.p, sizeof(int) * 3 * (size_t)m, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return n;
}

extern "C" long mcmb_lspace_weights(mcmb_ctx* c, double* w, long cap)
{
    if (!c || !c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_weights: no l-space setup");
    if (cap < 0 || (cap > 0 && !w)) return fail(MCMB_ERR_ARG, "mcmb_lspace_weights: bad buffer");
    MCMB_TRY(set_device(c));
    DevBuf rowptr, ws;
    MCMB_TRY(ensure_lspace_aux(c, c->ls, rowptr, ws, 0));
    long m = std::min(cap, c->ls.npoints);
    if (m > 0) MCMB_HIP(hipMemcpyAsync(w, ws.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return c->ls.npoints;
}
