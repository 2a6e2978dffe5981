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
constexpr int MNT = 512;         // threads (8 waves, two per SIMD)
constexpr int WN_T = 64;         // wave sub-tile columns
constexpr int NPIECE = 4 * 256 / MNT;   // 16-byte point chunks formed per thread and operand
constexpr int NB_C = WN_T / 32;         // 32-wide MFMA blocks per wave row
// waves tile the 256 x 256 mode tile as a (MT/128) x (MT/WN_T) grid of 128 x WN_T sub-tiles
static_assert((MNT / 64) == (MT / 128) * (MT / WN_T), "wave grid must cover the mode tile exactly");
static_assert(NPIECE * (MNT / MT) == 4, "threads per mode must cover the 4 point pieces");
static_assert(NPIECE == 2 && NB_C == 2, "the chunk schedule below interleaves 4 formation units with 4 MFMA groups");
// Chunks per segment = the FP32 accumulation span (8192 points).  The FP32
// accumulation error grows about linearly with the span: at lmax=2000, 1540
// modes the max elementwise Gram error is 1.1e-7 / 3.1e-7 / 2.4e-6 / 5.1e-6
// at 32 / 256 / 1024 / 2048 chunks, while the time is flat from 256 up and
// rises below it (more FP64 flushes).
constexpr int SEGCH_DEFAULT = 256;
constexpr int QP = 32;           // padded basis width

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

#define MCMB_SGB_1 __builtin_amdgcn_sched_group_barrier(0x002, 5, 0); __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
#define MCMB_SGB_VALU_MFMA MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1 \
                           MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1 MCMB_SGB_1
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16-byte chunk c (0..3) of operand row m, XOR-swizzled so that the rows one
// ds_read_b128 / ds_write_b128 pass touches land on distinct banks.
// Conflict-free both for 8-lane passes over 32 banks (rows m..m+7) and for
// 16-lane passes over 64 banks (rows m..m+15): the XOR key takes four
// distinct values over bits (1,2) of m for fixed bit 3, and over bits (2,3)
// for fixed bit 1.
__device__ __forceinline__ int swz(int m, int c)
{
    const int key = ((m >> 2) & 1) | ((((m >> 1) ^ (m >> 3)) & 1) << 1);
    return (m << 2) + (c ^ key);
}

// Row stride of the staged chunk basis: 36 floats puts the rows of eight
// consecutive third indices t on eight distinct 4-bank groups, so the
// ds_read_b128 of qT[t] by eight lanes with different t is conflict-free.
constexpr int QTS = MKC + 4;

struct MixedSmem {
    bf16x8 op[2][4][MT * 4];   // [buf][A_hi, A_lo, B_hi, B_lo][row*4 + chunk]
    float qT[2][QP][QTS];      // weighted chunk basis sqrt(w_kk)/6 q_c(k_kk), [c][kk]
    float stg[2][2 * 32];      // q(i)[32], q(j)[32]
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
        const float w = (float)ws[pbase + (tc - c0)];
        g.q[u] = (c < R && t <= last) ? w * v : 0.f;
    }
    // tid 0..31: q(i)[c]; 32..63: q(j)[c]
    const int c = tid & 31;
    const int l = (tid < 32) ? rd.i : rd.j;
    const float qv = qf[(long)l * QP + c];
    g.x = c < R ? qv : 0.f;
    return g;
}

__device__ __forceinline__ void stage_store(MixedSmem& S, int b, const StageRegs& g, int tid)
{
#pragma unroll
    for (int u = 0; u < 1024 / MNT; ++u) {
        const int e = tid + u * MNT;
        S.qT[b][e / MKC][e % MKC] = g.q[u];
    }
    if (tid < 64) S.stg[b][tid] = g.x;
}

// Per-chunk coefficients of one operand row: Q_m on the points of row (i,j)
// is Srs q_t(k) + Srt q_s(k) + Sst q_r(k) with S_ab = q_a(i) q_b(j) + q_b(i) q_a(j).
struct ModeCoef {
    float Srs, Srt, Sst;
    const float *qt, *qs, *qr;   // weighted chunk basis rows in LDS
};

__device__ __forceinline__ ModeCoef mode_coef(const MixedSmem& S, int b, int r, int s, int t, bool ok)
{
    const float* qi = &S.stg[b][0];
    const float* qj = &S.stg[b][32];
    const float ir = qi[r], is = qi[s], it = qi[t];
    const float jr = qj[r], js = qj[s], jt = qj[t];
    ModeCoef k;
    k.Srs = ok ? ir * js + is * jr : 0.f;
    k.Srt = ok ? ir * jt + it * jr : 0.f;
    k.Sst = ok ? is * jt + it * js : 0.f;
    k.qt = &S.qT[b][t][0];
    k.qs = &S.qT[b][s][0];
    k.qr = &S.qT[b][r][0];
    return k;
}

__device__ __forceinline__ unsigned pack_bf16x2(float x0, float x1)
{
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    const b2 h = __builtin_convertvector((f2){x0, x1}, b2);   // one v_cvt_pk_bf16_f32, RNE
    return __builtin_bit_cast(unsigned, h);
}

// Formation of points 8c .. 8c+7 of one operand row: the LDS reads (issued
// early, so that MFMAs cover their latency) and the arithmetic + stores.
struct PieceIn {
    f32x4 t[2], s[2], r[2];
};

__device__ __forceinline__ PieceIn piece_load(const ModeCoef& k, int c)
{
    PieceIn in;
#pragma unroll
    for (int q4 = 0; q4 < 2; ++q4) {
        in.t[q4] = reinterpret_cast<const f32x4*>(k.qt)[2 * c + q4];
        in.s[q4] = reinterpret_cast<const f32x4*>(k.qs)[2 * c + q4];
        in.r[q4] = reinterpret_cast<const f32x4*>(k.qr)[2 * c + q4];
    }
    return in;
}

// Writes the hi and lo parts of the 8 points into row m; returns sum lo^2
// over the 8 points (as rounded to BF16) when LO2.
template <bool LO2>
__device__ __forceinline__ float piece_store(const ModeCoef& k, const PieceIn& in, int c, int m,
                                             bf16x8* __restrict__ hi, bf16x8* __restrict__ lo)
{
    float x[8];
#pragma unroll
    for (int q4 = 0; q4 < 2; ++q4)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            x[q4 * 4 + e] = k.Srs * in.t[q4][e] + k.Srt * in.s[q4][e] + k.Sst * in.r[q4][e];
    unsigned hw[4], lw[4];
    float lo2 = 0.f;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const unsigned h = pack_bf16x2(x[2 * p], x[2 * p + 1]);
        const float h0 = __builtin_bit_cast(float, h << 16);
        const float h1 = __builtin_bit_cast(float, h & 0xffff0000u);
        const unsigned l = pack_bf16x2(x[2 * p] - h0, x[2 * p + 1] - h1);
        hw[p] = h;
        lw[p] = l;
        if (LO2) {
            const float l0 = __builtin_bit_cast(float, l << 16);
            const float l1 = __builtin_bit_cast(float, l & 0xffff0000u);
            lo2 += l0 * l0 + l1 * l1;
        }
    }
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    hi[swz(m, c)] = __builtin_bit_cast(bf16x8, (u4){hw[0], hw[1], hw[2], hw[3]});
    lo[swz(m, c)] = __builtin_bit_cast(bf16x8, (u4){lw[0], lw[1], lw[2], lw[3]});
    return lo2;
}

template <bool LO2>
__device__ __forceinline__ float form_piece(const ModeCoef& k, int c, int m, bf16x8* __restrict__ hi,
                                            bf16x8* __restrict__ lo)
{
    return piece_store<LO2>(k, piece_load(k, c), c, m, hi, lo);
}

template <bool DIAG>
__global__ __launch_bounds__(MNT, 1) void gram_tet_mixed_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr,
    const int2* __restrict__ seg_start, long seg0, long nchunks_total, int segch,
    const int2* __restrict__ tiles, int ntiles, const int* __restrict__ modes, int n,
    const float* __restrict__ qf, int R, int kstep, const double* __restrict__ ws,
    double* __restrict__ gacc, double* __restrict__ corr, int n_pad)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    MixedSmem& S = *reinterpret_cast<MixedSmem*>(smem_raw);
    const long seg = seg0 + blockIdx.x / ntiles;
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
        {
            const ModeCoef kA = mode_coef(S, 0, rA, sA, tA, okA);
#pragma unroll
            for (int u = 0; u < NPIECE; ++u) lo2A += form_piece<DIAG>(kA, c_lo + u, ml, S.op[0][0], S.op[0][1]);
            if (!DIAG) {
                const ModeCoef kB = mode_coef(S, 0, rB, sB, tB, okB);
#pragma unroll
                for (int u = 0; u < NPIECE; ++u) form_piece<false>(kB, c_lo + u, ml, S.op[0][2], S.op[0][3]);
            }
        }
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
            // Operand fragments of one K step: A rows wm*128 + a*32 + (lane&31),
            // B rows wn*64 + c*32 + (lane&31), 16-byte chunk ks*2 + lane/32.
            bf16x8 ah[4], al[4], bh[NB_C], bl[NB_C];
            auto ldA = [&](int ks, int a) {
                const int i = swz(wm * 128 + a * 32 + (lane & 31), ks * 2 + (lane >> 5));
                ah[a] = Ah[i];
                al[a] = Al[i];
            };
            auto ldB = [&](int ks, int c) {
                const int i = swz(wn * WN_T + c * 32 + (lane & 31), ks * 2 + (lane >> 5));
                bh[c] = Bh[i];
                bl[c] = Bl[i];
            };
            auto mfma3 = [&](int a, int c) {
                acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bh[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[a], bl[c], acc[a][c], 0, 0, 0);
                acc[a][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[a], bh[c], acc[a][c], 0, 0, 0);
            };
#pragma unroll
            for (int a = 0; a < 4; ++a) ldA(0, a);
#pragma unroll
            for (int c = 0; c < NB_C; ++c) ldB(0, c);
            // MFMAs on chunk ch in four groups (K step ks, column block c),
            // each paired with one formation unit of chunk ch+1 into the other
            // buffer: the unit's LDS reads go first, the MFMAs cover their
            // latency, then its arithmetic and stores, which the scheduler
            // interleaves with the MFMAs in flight.  Units: A pieces then B
            // pieces (off-diagonal); A pieces in the even units (diagonal).
            // The fragments of K step 1 are loaded in group (0, 1), each
            // right after the last use of the registers it replaces.  After
            // the last chunk the formed data is unused.
            const ModeCoef kA = mode_coef(S, b ^ 1, rA, sA, tA, okA);
            const ModeCoef kB = mode_coef(S, b ^ 1, rB, sB, tB, okB);
            float l2 = 0.f;
#pragma unroll
            for (int unit = 0; unit < 4; ++unit) {
                const int c = unit & 1;
                const bool formA = DIAG ? (unit & 1) == 0 : unit < NPIECE;
                const bool formB = !DIAG && unit >= NPIECE;
                const int piece = c_lo + (DIAG ? unit / 2 : (unit & 1));
                PieceIn in;
                if (formA) in = piece_load(kA, piece);
                if (formB) in = piece_load(kB, piece);
                if (unit == 1) ldB(1, 0);
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    mfma3(a, c);
                    if (unit == 1) ldA(1, a);
                }
                if (unit == 1) ldB(1, 1);
                if (formA) l2 += piece_store<DIAG>(kA, in, piece, ml, S.op[b ^ 1][0], S.op[b ^ 1][1]);
                if (formB) piece_store<false>(kB, in, piece, ml, S.op[b ^ 1][2], S.op[b ^ 1][3]);
#ifndef MCMB_NO_SGB
                // order within the group: LDS reads, two MFMAs, then one MFMA
                // per five VALU instructions
                __builtin_amdgcn_sched_group_barrier(0x100, 6, 0);
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
                MCMB_SGB_VALU_MFMA __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
#endif
                __builtin_amdgcn_sched_barrier(0);
            }
            lo2A += (ch + 1 < nch) ? l2 : 0.f;

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
    // segment) at the price of longer FP32 accumulation; MCMB_MIXED_SEGCH
    // overrides the length.
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
    // one dispatch holds fewer than 2^32 work-items: launch the segments in
    // batches of at most max_wg / ntiles segments (MCMB_MIXED_MAX_WG lowers
    // the cap, so that tests reach the batching at small sizes)
    long max_wg = ((1ll << 32) - 1) / MNT;
    if (const char* e = getenv("MCMB_MIXED_MAX_WG")) max_wg = std::max(1l, std::min(max_wg, atol(e)));
    int launches = 0;
    auto launch_all = [&](bool diag, const std::vector<int2>& tl, const int2* d_tl) -> int {
        const long nt = (long)tl.size();
        const long per = std::max(1l, max_wg / nt);
        for (long s0 = 0; s0 < nseg; s0 += per) {
            const long ns = std::min(per, nseg - s0);
            MCMB_TRY(launch_guard("mixed Gram", (unsigned long long)(ns * nt), MNT));
            if (diag)
                hipLaunchKernelGGL(gram_tet_mixed_kernel<true>, dim3((unsigned)(ns * nt)), dim3(MNT), smem, c->stream,
                                   L.d_rows.as<RowDesc>(), d_rowptr, d_seg.as<int2>(), s0, nchunks, segch, d_tl,
                                   (int)nt, L.d_modes.as<int>(), n, L.d_qf.as<float>(), R, L.kstep, d_ws,
                                   c->d_partial.as<double>(), d_corr.as<double>(), n_pad);
            else
                hipLaunchKernelGGL(gram_tet_mixed_kernel<false>, dim3((unsigned)(ns * nt)), dim3(MNT), smem, c->stream,
                                   L.d_rows.as<RowDesc>(), d_rowptr, d_seg.as<int2>(), s0, nchunks, segch, d_tl,
                                   (int)nt, L.d_modes.as<int>(), n, L.d_qf.as<float>(), R, L.kstep, d_ws,
                                   c->d_partial.as<double>(), d_corr.as<double>(), n_pad);
            MCMB_HIP(hipGetLastError());
            ++launches;
        }
        return MCMB_OK;
    };
    if (!otiles.empty()) MCMB_TRY(launch_all(false, otiles, d_ot.as<int2>()));
    MCMB_TRY(launch_all(true, dtiles, d_dt.as<int2>()));
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

