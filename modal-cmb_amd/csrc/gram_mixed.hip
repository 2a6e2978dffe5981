// Mixed-precision tetrapyd Gram on BF16 MFMA (gfx950 v_mfma_f32_32x32x16_bf16).
//
// Same decomposition as the FP64 path (rows of the ordered tetrapyd cut into
// chunks that never cross a row; triangle of mode tiles x point segments), but
// the operand a = sqrt(w) Q_m is formed in FP32 and split into two BF16 parts,
// a = hi + lo + O(2^-17 |a|).  Each tile accumulates
//   hi_A hi_B + hi_A lo_B + lo_A hi_B
// in FP32 MFMA accumulators over one segment of SEGCH chunks (any rows), then
// added into the FP64 accumulator of the whole Gram with no-return FP64
// atomics (the order of segment sums is not fixed: results agree to FP64
// rounding of the segment sums, far below the split error).  The
// dropped lo_A lo_B product is a random-sign term off the diagonal but a
// positive bias sum_p lo_pm^2 on it, so diagonal tiles also sum lo^2 per mode
// in FP64 (VALU) and the reduction adds it to Gamma_mm.
//
// Workgroup: 8 waves (two per SIMD, so one wave's operand formation overlaps
// the other's MFMA work), 2 x 4 grid of 128 x 64 wave sub-tiles of a 256 x 256
// mode tile; thread t forms 16 of the 32 points of mode t % 256 of each operand.  Diagonal and off-diagonal tiles are
// separate instantiations (the diagonal one forms a single operand).
#include "mcmb_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace mcmb {

constexpr int MT = 256;          // mode tile
constexpr int MKC = 32;          // points per chunk (two K=16 MFMA steps)
#ifdef MCMB_MIXED_4W
constexpr int MNT = 256;         // threads (4 waves, one per SIMD)
constexpr int WN_T = 128;        // wave sub-tile columns
#else
constexpr int MNT = 512;         // threads (8 waves, two per SIMD)
constexpr int WN_T = 64;         // wave sub-tile columns
#endif
constexpr int NPIECE = 4 * 256 / MNT;   // 16-byte point chunks formed per thread and operand
constexpr int NB_C = WN_T / 32;         // 32-wide MFMA blocks per wave row
// waves tile the 256 x 256 mode tile as a (MT/128) x (MT/WN_T) grid of 128 x WN_T sub-tiles
static_assert((MNT / 64) == (MT / 128) * (MT / WN_T), "wave grid must cover the mode tile exactly");
static_assert(NPIECE * (MNT / MT) == 4, "threads per mode must cover the 4 point pieces");
constexpr int SEGCH_DEFAULT = 2048;  // chunks per segment (FP32 accumulation span)
constexpr int QP = 32;           // padded basis width

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16-byte chunk c (0..3) of operand row m, XOR-swizzled so that the 8 rows a
// ds_read_b128 phase touches land on distinct banks.
__device__ __forceinline__ int swz(int m, int c) { return (m << 2) + (c ^ ((m >> 1) & 3)); }

struct MixedSmem {
    bf16x8 op[2][4][MT * 4];   // [buf][A_hi, A_lo, B_hi, B_lo][row*4 + chunk]
    float qT[2][QP][MKC];      // chunk basis q_c(k_kk), [c][kk]
    float stg[2][3 * 32];      // q(i)[32], q(j)[32], sqrt(w)/6 [32]
};

// Global-load half of the staging of chunk (rd, c0); branch-free (addresses
// clamped into the row, invalid entries selected to zero).
struct StageRegs {
    float q[1024 / MNT], x;
};

__device__ __forceinline__ StageRegs stage_load(const RowDesc& rd, int c0, long pbase, int kstep,
                                                const float* __restrict__ qf, int R,
                                                const double* __restrict__ ws, int tid)
{
    StageRegs g;
    const int last = rd.nk - 1;
#pragma unroll
    for (int u = 0; u < 1024 / MNT; ++u) {
        const int e = tid + u * MNT, c = e / MKC, kk = e % MKC, t = c0 + kk;
        const int tc = t < last ? t : last;
        const float v = qf[(long)(rd.k0 + kstep * tc) * QP + c];
        g.q[u] = (c < R && t <= last) ? v : 0.f;
    }
    // tid 0..31: q(i)[c]; 32..63: q(j)[c]; 64..95: w of point c0 + c
    const int c = tid & 31;
    const int l = (tid < 32) ? rd.i : rd.j;
    const float qv = qf[(long)l * QP + c];
    const int t = c0 + c;
    const double wv = ws[pbase + (t <= last ? c : last - c0)];
    g.x = (tid < 64) ? (c < R ? qv : 0.f) : ((t <= last) ? (float)wv : 0.f);
    return g;
}

__device__ __forceinline__ void stage_store(MixedSmem& S, int b, const StageRegs& g, int tid)
{
#pragma unroll
    for (int u = 0; u < 1024 / MNT; ++u) (&S.qT[b][0][0])[tid + u * MNT] = g.q[u];
    if (tid < 96) S.stg[b][tid] = g.x;
}

// Form points 8 c_lo .. 8 (c_lo + NPIECE) - 1 of mode (r,s,t) into hi/lo row m;
// returns sum lo^2.
__device__ __forceinline__ float form_mixed(const MixedSmem& S, int b, int r, int s, int t, bool ok, int m,
                                            int c_lo, bf16x8* __restrict__ hi, bf16x8* __restrict__ lo)
{
    const float* qi = &S.stg[b][0];
    const float* qj = &S.stg[b][32];
    float ir = qi[r], is = qi[s], it = qi[t];
    float jr = qj[r], js = qj[s], jt = qj[t];
    float Srs = ir * js + is * jr;
    float Srt = ir * jt + it * jr;
    float Sst = is * jt + it * js;
    if (!ok) Srs = Srt = Sst = 0.f;
    const f32x4* qt = reinterpret_cast<const f32x4*>(&S.qT[b][t][0]);
    const f32x4* qs = reinterpret_cast<const f32x4*>(&S.qT[b][s][0]);
    const f32x4* qr = reinterpret_cast<const f32x4*>(&S.qT[b][r][0]);
    const f32x4* wv = reinterpret_cast<const f32x4*>(&S.stg[b][64]);
    float lo2 = 0.f;
#pragma unroll
    for (int c = c_lo; c < c_lo + NPIECE; ++c) {   // 16-byte chunk c = points 8c..8c+7
        // bound the LDS reads the scheduler hoists (each piece is 32 floats)
        __builtin_amdgcn_sched_barrier(0);
        bf16x8 h8, l8;
#pragma unroll
        for (int q4 = 0; q4 < 2; ++q4) {
            const f32x4 a = qt[c * 2 + q4], bq = qs[c * 2 + q4], cq = qr[c * 2 + q4], w4 = wv[c * 2 + q4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float x = w4[e] * (Srs * a[e] + Srt * bq[e] + Sst * cq[e]);
                const __bf16 xh = (__bf16)x;
                const __bf16 xl = (__bf16)(x - (float)xh);
                const float fl = (float)xl;
                lo2 += fl * fl;
                h8[q4 * 4 + e] = xh;
                l8[q4 * 4 + e] = xl;
            }
        }
        hi[swz(m, c)] = h8;
        lo[swz(m, c)] = l8;
    }
    return lo2;
}

template <bool DIAG>
__global__ __launch_bounds__(MNT, 1) void gram_tet_mixed_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr,
    const int2* __restrict__ seg_start, long nchunks_total, int segch,
    const int2* __restrict__ tiles, int ntiles, const int* __restrict__ modes, int n,
    const float* __restrict__ qf, int R, int kstep, const double* __restrict__ ws,
    double* __restrict__ gacc, double* __restrict__ corr, int n_pad)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    MixedSmem& S = *reinterpret_cast<MixedSmem*>(smem_raw);
    const long seg = blockIdx.x / ntiles;
    const int tile = blockIdx.x % ntiles;
    const int2 IJ = tiles[tile];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / (MT / WN_T) , wn = wave % (MT / WN_T);
    const int ml = tid & (MT - 1), c_lo = (tid >> 8) * NPIECE;

    const int mAg = IJ.x * MT + ml, mBg = IJ.y * MT + ml;
    const bool okA = mAg < n, okB = mBg < n;
    const int pkA = okA ? modes[mAg] : 0, pkB = okB ? modes[mBg] : 0;
    const int rA = pkA & 255, sA = (pkA >> 8) & 255, tA = (pkA >> 16) & 255;
    const int rB = pkB & 255, sB = (pkB >> 8) & 255, tB = (pkB >> 16) & 255;

    f32x16 acc[4][NB_C];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < NB_C; ++b)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[a][b][v] = 0.f;

    // this wave's corner of the FP64 Gram accumulator
    double* const Pw = gacc + (size_t)(IJ.x * MT + wm * 128 + 4 * (lane >> 5)) * n_pad + IJ.y * MT + wn * WN_T + (lane & 31);
    auto flush = [&]() {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int c = 0; c < NB_C; ++c) {
                double* pb = Pw + (size_t)(a * 32) * n_pad + c * 32;
#pragma unroll
                for (int v = 0; v < 16; ++v)
                    unsafeAtomicAdd(pb + (size_t)(8 * (v >> 2) + (v & 3)) * n_pad, (double)acc[a][c][v]);
            }
    };

    double lo2A = 0.0;
    const long rest = nchunks_total - seg * segch;
    const int nch = rest < segch ? (int)rest : segch;
    // chunk position walker (uniform over the workgroup)
    int prow = seg_start[seg].x, pc0 = seg_start[seg].y;   // next chunk to stage
    auto stage_next = [&](StageRegs& g) {
        const RowDesc rd = rows[prow];
        g = stage_load(rd, pc0, rowptr[prow] + pc0, kstep, qf, R, ws, tid);
        pc0 += MKC;
        if (pc0 >= rd.nk) { ++prow; pc0 = 0; }
    };

    if (nch > 0) {
        StageRegs g;
        stage_next(g);
        stage_store(S, 0, g, tid);
        __syncthreads();
        lo2A += form_mixed(S, 0, rA, sA, tA, okA, ml, c_lo, S.op[0][0], S.op[0][1]);
        if (!DIAG) form_mixed(S, 0, rB, sB, tB, okB, ml, c_lo, S.op[0][2], S.op[0][3]);
        if (nch > 1) {
            stage_next(g);
            stage_store(S, 1, g, tid);
        }
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
            const bool more2 = ch + 2 < nch;
            if (more2) stage_next(g);
            const bf16x8* Ah = S.op[b][0];
            const bf16x8* Al = S.op[b][1];
            const bf16x8* Bh = DIAG ? S.op[b][0] : S.op[b][2];
            const bf16x8* Bl = DIAG ? S.op[b][1] : S.op[b][3];
#pragma unroll
            for (int ks = 0; ks < 2; ++ks) {
                // all 8 operand fragments of this K step once (16 ds_read_b128),
                // then 48 MFMAs
                const int cidx = ks * 2 + (lane >> 5);
                bf16x8 ah[4], al[4], bh[NB_C], bl[NB_C];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int ma = wm * 128 + a * 32 + (lane & 31);
                    ah[a] = Ah[swz(ma, cidx)];
                    al[a] = Al[swz(ma, cidx)];
                }
#pragma unroll
                for (int c = 0; c < NB_C; ++c) {
                    const int mb = wn * WN_T + c * 32 + (lane & 31);
                    bh[c] = Bh[swz(mb, cidx)];
                    bl[c] = Bl[swz(mb, cidx)];
                }
#pragma unroll
                for (int c = 0; c < NB_C; ++c)
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
                        acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
                        acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
                    }
            }
            // form chunk ch+1 (unconditionally, so that it shares the basic block
            // with the MFMAs above; after the last chunk the result is unused)
            {
                const float l2 = form_mixed(S, b ^ 1, rA, sA, tA, okA, ml, c_lo, S.op[b ^ 1][0], S.op[b ^ 1][1]);
                lo2A += (ch + 1 < nch) ? l2 : 0.f;
                if (!DIAG) form_mixed(S, b ^ 1, rB, sB, tB, okB, ml, c_lo, S.op[b ^ 1][2], S.op[b ^ 1][3]);
            }

            if (more2) stage_store(S, b, g, tid);
            __syncthreads();
        }
    }
    flush();
    if (DIAG && okA) unsafeAtomicAdd(corr + mAg, lo2A);   // both point halves add
}

// Gamma[m][c] = Gamma[c][m] = acc[m][c] (m <= c), + corr[m] on the diagonal.
__global__ __launch_bounds__(256) void gram_mixed_finalize_kernel(const double* __restrict__ gacc,
                                                                  const double* __restrict__ corr, int n_pad,
                                                                  int n, double* __restrict__ out)
{
    int c = blockIdx.x * 64 + (threadIdx.x & 63);
    int m = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (m >= n || c >= n || m > c) return;
    double s = gacc[(size_t)m * n_pad + c];
    if (m == c) s += corr[m];
    out[(size_t)m * n + c] = s;
    out[(size_t)c * n + m] = s;
}

__global__ void to_float_table(const double* __restrict__ q, int ldq, int R, int nrow, float* __restrict__ qf)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nrow * QP) return;
    int l = e / QP, c = e % QP;
    qf[e] = c < R ? (float)q[(long)l * ldq + c] : 0.f;
}

int gram_mixed_launch(mcmb_ctx* c, PointSet& L, const long* d_rowptr, const double* d_ws, double* d_gamma,
                      int row_lo, int row_hi)
{
    const int n = L.n, R = L.R;
    if (R > QP) return fail(MCMB_ERR_ARG, "mixed Gram supports basis width <= %d (got %d)", QP, R);
    const int T = (n + MT - 1) / MT, n_pad = T * MT;
    std::vector<int2> dtiles, otiles;
    for (int I = 0; I < T; ++I)
        for (int J = I; J < T; ++J) (I == J ? dtiles : otiles).push_back(make_int2(I, J));
    const long ntiles = (long)T * (T + 1) / 2;
    // segments: segch consecutive chunks in walk order (start position kept).
    // Longer segments cut the FP64 flush traffic (n_pad^2/2 doubles per
    // segment) at the price of longer FP32 accumulation.
    long nchunks_all = 0;
    for (int r = row_lo; r < row_hi; ++r) nchunks_all += (L.rows[r].nk + MKC - 1) / MKC;
    // keep at least ~4 rounds of 256 workgroups on small problems
    long cap = std::max<long>(16, nchunks_all * ntiles / 1024);
    int segch = (int)std::min<long>(SEGCH_DEFAULT, cap);
    if (const char* e = getenv("MCMB_MIXED_SEGCH")) segch = std::max(1, atoi(e));
    std::vector<int2> seg;
    long nchunks = 0;
    for (int r = row_lo; r < row_hi; ++r)
        for (int c0 = 0; c0 < L.rows[r].nk; c0 += MKC) {
            if (nchunks % segch == 0) seg.push_back(make_int2(r, c0));
            ++nchunks;
        }
    if (nchunks == 0) {   // empty slice: a zero Gram
        MCMB_HIP(hipMemsetAsync(d_gamma, 0, sizeof(double) * (size_t)n * n, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        c->last_ms = 0.0;
        c->last_launches = 0;
        return MCMB_OK;
    }
    const long nseg = (long)seg.size();

    DevBuf d_dt, d_ot, d_seg, d_corr;
    MCMB_TRY(d_dt.alloc(sizeof(int2) * dtiles.size()));
    if (!otiles.empty()) MCMB_TRY(d_ot.alloc(sizeof(int2) * otiles.size()));
    MCMB_TRY(d_seg.alloc(sizeof(int2) * seg.size()));
    MCMB_TRY(d_corr.alloc(sizeof(double) * (size_t)n_pad));
    const size_t abytes = (size_t)n_pad * n_pad * sizeof(double);
    MCMB_TRY(c->d_partial.alloc(abytes));
    MCMB_TRY(L.d_qf.alloc(sizeof(float) * (size_t)(L.lmax + 1) * QP));
    MCMB_HIP(hipMemcpyAsync(d_dt.p, dtiles.data(), sizeof(int2) * dtiles.size(), hipMemcpyHostToDevice, c->stream));
    if (!otiles.empty())
        MCMB_HIP(hipMemcpyAsync(d_ot.p, otiles.data(), sizeof(int2) * otiles.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_seg.p, seg.data(), sizeof(int2) * seg.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemsetAsync(d_corr.p, 0, sizeof(double) * (size_t)n_pad, c->stream));
    MCMB_HIP(hipMemsetAsync(c->d_partial.p, 0, abytes, c->stream));
    int ne = (L.lmax + 1) * QP;
    hipLaunchKernelGGL(to_float_table, dim3((ne + 255) / 256), dim3(256), 0, c->stream, L.d_q.as<double>(), L.ldq,
                       R, L.lmax + 1, L.d_qf.as<float>());
    MCMB_HIP(hipGetLastError());
    const size_t smem = sizeof(MixedSmem);
    MCMB_HIP(hipFuncSetAttribute((const void*)gram_tet_mixed_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    MCMB_HIP(hipFuncSetAttribute((const void*)gram_tet_mixed_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    int launches = 0;
    if (!otiles.empty()) {
        hipLaunchKernelGGL(gram_tet_mixed_kernel<false>, dim3((unsigned)(otiles.size() * nseg)), dim3(MNT), smem,
                           c->stream, L.d_rows.as<RowDesc>(), d_rowptr, d_seg.as<int2>(), nchunks, segch,
                           d_ot.as<int2>(), (int)otiles.size(), L.d_modes.as<int>(), n, L.d_qf.as<float>(), R, L.kstep,
                           d_ws, c->d_partial.as<double>(), d_corr.as<double>(), n_pad);
        MCMB_HIP(hipGetLastError());
        ++launches;
    }
    hipLaunchKernelGGL(gram_tet_mixed_kernel<true>, dim3((unsigned)(dtiles.size() * nseg)), dim3(MNT), smem,
                       c->stream, L.d_rows.as<RowDesc>(), d_rowptr, d_seg.as<int2>(), nchunks, segch,
                       d_dt.as<int2>(), (int)dtiles.size(), L.d_modes.as<int>(), n, L.d_qf.as<float>(), R, L.kstep,
                       d_ws, c->d_partial.as<double>(), d_corr.as<double>(), n_pad);
    MCMB_HIP(hipGetLastError());
    ++launches;
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    dim3 rgrid((unsigned)((n + 63) / 64), (unsigned)((n + 3) / 4));
    hipLaunchKernelGGL(gram_mixed_finalize_kernel, rgrid, dim3(256), 0, c->stream, c->d_partial.as<double>(),
                       d_corr.as<double>(), n_pad, n, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = launches;
    return MCMB_OK;
}

}  // namespace mcmb

