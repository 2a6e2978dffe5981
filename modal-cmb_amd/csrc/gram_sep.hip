// Separable tetrapyd Gram: Gamma_mn = sum_{(l1,l2,l3) in T} w(l) Q_m(l) Q_n(l)
// through three successive contractions of the pair products of the basis.
//
// With Q_m(l) = (1/6) sum_pi prod_a q_{m_pi(a)}(l_a) ((m_1, m_2, m_3) = order[m]),
//   Q_m(l) Q_n(l) = (1/36) sum_{pi, rho} prod_a U_{m_pi(a) n_rho(a)}(l_a),
//   U_{xy}(l) = q_x(l) q_y(l)  (indexed by the unordered pair, NP = R(R+1)/2),
// so Gamma_mn = (1/36) sum_{pi, rho} G[{m_pi1 n_rho1}][{m_pi2 n_rho2}][{m_pi3 n_rho3}]
// with the slot-ordered tensor
//   G[a][e][f] = sum_{l in T} w(l) U_a(l1) U_e(l2) U_f(l3),
// formed in three stages over the rows (l1, l2) of the ordered tetrapyd:
//   1. M[row][f]    = sum_{l3 in row} w U_f(l3)          (row moments, FP64 MFMA)
//   2. Y[i][e][f]   = sum_{rows of l1 = i} U_e(l2) M[row][f]   (FP64 MFMA, batched per i)
//   3. G2[a][{e,f}] = sum_i U_a(i) (Y[i][e][f] + Y[i][f][e])   (FP64 MFMA, one GEMM
//                     over the NP(NP+1)/2 unordered pairs {e, f})
// Only the fully symmetrised tensor enters Gamma: summing the 36 slot
// assignments over the 6 orders of the three matched pairs first,
//   Gamma_mn = (1/36) sum_{sigma in S3} Gs(p_{1 s1}, p_{2 s2}, p_{3 s3}),
//   p_ab = {m_a, n_b},  Gs(x, y, z) = G2[x][{y,z}] + G2[y][{x,z}] + G2[z][{x,y}]
// (Gs = the sum of G over the 6 orders of its slots, kept for x <= y <= z),
// so stage 3 runs over half the columns and the assembly gathers 6 entries of
// a 0.14 GB table.  At lmax 2500 with R = 30 (NP = 465) the stages cost
// 3.0e11, 6.7e11 and 1.3e11 multiply-adds, against 8e15 for the
// point-by-point contraction; the summation is FP64 throughout and in a
// fixed order (no atomics), so repeated calls are bit-identical.
#include "mcmb_internal.h"
#include "weights.h"
#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

namespace mcmb {

typedef double fd4 __attribute__((ext_vector_type(4)));
typedef double fd2 __attribute__((ext_vector_type(2)));

constexpr int SG_FM = 5, SG_FN = 3;   // 16 x 16 blocks per wave (rows, columns); 2 x 2 waves
constexpr int SG_TM = 32 * SG_FM;     // GEMM tile rows (160)
constexpr int SG_TN = 32 * SG_FN;     // GEMM tile columns (96)
constexpr int SG_SMALL = 64;          // basis widths up to this use 64 x 64 tiles (2 x 2 blocks per wave)
constexpr int SG_KC = 16;             // K rows per LDS stage
constexpr int SG_KSPLIT = 16;         // K splits of the l1 contraction over a compressed basis
constexpr int SR_UPAD = 8;          // zero rows after the last l of the basis tables
constexpr int SR_G = 16;            // rows per row-moment group
constexpr int SR_KC = 16;           // l3 values per weight stage (SR_G * SR_KC = 256 threads)
constexpr int SR_NPMAX = 768;       // pair columns of stage 1 (3 per thread of 256)

struct SepTask {
    long boff;   // first B row of the task
    long coff;   // first C element of the task
    int aoff;    // first entry of the A row index list
    int K;       // rows of the task
};

// U[l][f] = q_x(l) q_y(l), f = pair (x, y)
// (padding columns f >= NP are zero)
__global__ void sep_utab(const double* __restrict__ q, int ldq, int nl, const int2* __restrict__ pxy, int NP, int ld,
                         double* __restrict__ U)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(nl + SR_UPAD) * ld) return;
    const int l = (int)(e / ld), f = (int)(e % ld);
    if (f >= NP || l >= nl) {
        U[e] = 0.0;
        return;
    }
    const int2 xy = pxy[f];
    U[e] = q[(long)l * ldq + xy.x] * q[(long)l * ldq + xy.y];
}

// Weights of the row groups, staged densely for stage 1: for group gb of G
// rows, l3 index kg of the group's grid (l3 = kmin + kstep kg) and member row
// g, Wd[wofs[gb] + G kg + g] = w (0 outside the row's range), kg padded to a
// multiple of 16.  KIND 0: the l-space weight with the exact h^2 less its l3
// factor (sep_weight: the contraction's basis table carries (2 l3 + 1) v_l3);
// KIND 1: the whole weight permsix h^2 v v v with the asymptotic h^2; both
// formed here; KIND 2: copied from the walk-order weights ws.
template <int KIND, int G>
__global__ __launch_bounds__(256) void sep_wstage(const RowDesc* __restrict__ rows, const long* __restrict__ rowptr,
                                                  const double* __restrict__ ws, const int* __restrict__ groups,
                                                  const int4* __restrict__ ginfo, const long* __restrict__ wofs,
                                                  int kstep, const double* __restrict__ f,
                                                  const double* __restrict__ v, double* __restrict__ Wd)
{
    static_assert(256 % G == 0 && (256 / G) <= 16, "whole l3 values per pass, passes within a 16-value chunk");
    constexpr int KP = 256 / G;   // l3 values per pass
    __shared__ int gi[G], gj[G], gk0[G], gnk[G];
    __shared__ long gp[G];
    const int tid = threadIdx.x;
    if (tid < G) {
        const int r = groups[(long)blockIdx.x * G + tid];
        RowDesc d = {0, 0, 0, 0};
        long p = 0;
        if (r >= 0) {
            d = rows[r];
            if (KIND == 2) p = rowptr[r];
        }
        gi[tid] = d.i;
        gj[tid] = d.j;
        gk0[tid] = d.k0;
        gnk[tid] = d.nk;
        gp[tid] = p;
    }
    __syncthreads();
    const int4 info = ginfo[blockIdx.x];
    const int kmin = info.x, nkk = info.y, nkp = (nkk + 15) / 16 * 16;
    double* const W = Wd + wofs[blockIdx.x];
    const int kk = tid / G, g = tid % G;
    for (int k0 = 0; k0 < nkp; k0 += KP) {
        const int kg = k0 + kk;
        double w = 0.0;
        if (kg < nkk) {
            const int k = kmin + kstep * kg;
            const int d = k - gk0[g];
            if (d >= 0 && d % kstep == 0 && d / kstep < gnk[g]) {
                if (KIND == 2) {
                    w = ws[gp[g] + d / kstep];
                } else if (KIND == 1) {
                    w = lspace_weight(gi[g], gj[g], k, 1, f, v);
                } else {
                    const int i = gi[g], j = gj[g], gg = (i + j + k) >> 1;
                    w = sep_weight(sep_row_factor(i, j, permsix_d(i, j, k), v), f[2 * (gg - i)], f[2 * (gg - j)],
                                   f[2 * (gg - k)], f[2 * gg + 1]);
                }
            }
        }
        W[G * k0 + tid] = w;   // = W[G kg + g]
    }
}

// Stage 1.  One workgroup per group of up to 16 rows that share l1 and the
// parity class of l3, so that their l3 ranges interleave on one grid of step
// kstep.  The NT threads split into GS row sets of 16/GS rows; within a set
// thread t owns pair columns t, t + NT/GS, ... (F of them).  Per stage of 16
// l3 values the pair products U[l3] go to registers first (all loads in
// flight together) and the 16 x 16 staged weights to LDS (one coalesced load,
// double-buffered); every weight read from LDS (a broadcast) feeds F FMAs.
// H = parts of the stage, to bound the registers held for U.
template <int NT, int GS, int F, int H>
__global__ __launch_bounds__(NT) void sep_rowmom(const int* __restrict__ groups, const int4* __restrict__ ginfo,
                                                 const long* __restrict__ wofs, const double* __restrict__ Wd,
                                                 int row_base, int kstep, const double* __restrict__ U, int ldu,
                                                 int NP, int ld, double* __restrict__ M)
{
    constexpr int KH = SR_KC / H;
    constexpr int WPT = SR_G * SR_KC / NT;   // staged weights loaded per thread per stage
    constexpr int RG = SR_G / GS;            // rows per thread
    constexpr int CT = NT / GS;              // threads per row set
    __shared__ double Wt[2][SR_KC][SR_G];
    const int tid = threadIdx.x, ct = tid % CT, g0 = RG * (tid / CT);
    const int4 info = ginfo[blockIdx.x];
    const int kmin = info.x, nkk = info.y;
    const double* const Wg = Wd + wofs[blockIdx.x];

    double acc[RG][F];
#pragma unroll
    for (int g = 0; g < RG; ++g)
#pragma unroll
        for (int u = 0; u < F; ++u) acc[g][u] = 0.0;

    int buf = 0;
    for (int c0 = 0; c0 < nkk; c0 += SR_KC, buf ^= 1) {
        double wr[WPT];
#pragma unroll
        for (int x = 0; x < WPT; ++x) wr[x] = Wg[16 * c0 + WPT * tid + x];
#pragma unroll
        for (int h = 0; h < H; ++h) {
            double uu[KH][F];
            // l3 values past the group's range repeat the last one (weight 0)
#pragma unroll
            for (int kk = 0; kk < KH; ++kk) {
                const int kc = min(c0 + h * KH + kk, nkk - 1);
                const double* Uk = U + (long)(kmin + kstep * kc) * ldu;
#pragma unroll
                for (int u = 0; u < F; ++u) {
                    const int f = ct + CT * u;
                    uu[kk][u] = f < NP ? Uk[f] : 0.0;
                }
            }
            if (h == 0) {
#pragma unroll
                for (int x = 0; x < WPT; ++x) (&Wt[buf][0][0])[WPT * tid + x] = wr[x];
                __syncthreads();
            }
#pragma unroll
            for (int kk = 0; kk < KH; ++kk)
#pragma unroll
                for (int g = 0; g < RG; ++g) {
                    const double w = Wt[buf][h * KH + kk][g0 + g];
#pragma unroll
                    for (int u = 0; u < F; ++u) acc[g][u] = fma(w, uu[kk][u], acc[g][u]);
                }
        }
    }
#pragma unroll
    for (int g = 0; g < RG; ++g) {
        const int r = groups[(long)blockIdx.x * SR_G + g0 + g];
        if (r < 0) continue;
        double* Mr = M + (long)(r - row_base) * ld;
#pragma unroll
        for (int u = 0; u < F; ++u) {
            const int f = ct + CT * u;
            if (f < NP) Mr[f] = acc[g][u];
        }
    }
}

// Stage 1 on FP64 MFMA: one workgroup (blockDim.x / 64 waves) per group of
// 16 RT rows that share l1 and the parity class of l3.  M[row][f] =
// sum_kg W[row][kg] U[l3(kg)][f] as 16 x 16 x 4 products with A = the staged
// weights (rows x l3) and B = the pair products (l3 x columns).  Wave w owns
// the 16 RT rows x the 16 CPW columns from 16 CPW w (ldu >= 16 CPW x waves,
// padding columns of U are zero); every U element is loaded once per workgroup,
// straight into B registers, and feeds the RT row tiles.  The operands of
// K-steps s + 1 and s + 2 are in flight while step s computes.
constexpr int SR_GM = 32;   // rows per group of the MFMA kernel (RT = 2)
constexpr int SR_XC = 64;   // consecutive groups per XCD turn (fused l3 kernel)
template <int CPW, int RT>
__global__ __launch_bounds__(512, 2) void sep_rowmom_mfma(const int* __restrict__ groups,
                                                          const int4* __restrict__ ginfo,
                                                          const long* __restrict__ wofs,
                                                          const double* __restrict__ Wd, int row_base, int kstep,
                                                          const double* __restrict__ U, int ldu, int NP, int ld,
                                                          double* __restrict__ M)
{
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int4 info = ginfo[blockIdx.x];
    const int kmin = info.x, nkk = info.y;
    const int nks = (nkk + 3) >> 2;   // K-steps of 4 l3 values (the staged weights are zero-padded past nkk)
    constexpr int G = 16 * RT;
    const double* const Wg = Wd + wofs[blockIdx.x] + G * (lane >> 4) + (lane & 15);
    const int col0 = 16 * CPW * w + (lane & 15);
    const double* const Ub = U + (long)kmin * ldu + col0;
    const long ustep = (long)kstep * ldu;

    fd4 acc[RT][CPW];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CPW; ++ct) acc[rt][ct] = (fd4){0.0, 0.0, 0.0, 0.0};

    // K-step s: A = W[16 rt + lane % 16][4 s + lane / 16],
    //           B = U[l3(4 s + lane / 16)][col0 + 16 ct] (l3 index clamped to
    //           the last one, where the weights are zero)
    auto load = [&](int s, double (&a)[RT], double (&b)[CPW]) {
        const int kg = 4 * s + (lane >> 4);
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) a[rt] = Wg[G * 4 * s + 16 * rt];
        const double* Ur = Ub + ustep * min(kg, nkk - 1);
#pragma unroll
        for (int ct = 0; ct < CPW; ++ct) b[ct] = Ur[16 * ct];
    };
    auto step = [&](const double (&a)[RT], const double (&b)[CPW]) {
#pragma unroll
        for (int ct = 0; ct < CPW; ++ct)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
                acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
    };
    // three operand sets in rotation; reloads past the last K-step are clamped
    // to it (unused), so the loop is one straight-line body
    double a0[RT], b0[CPW], a1[RT], b1[CPW], a2[RT], b2[CPW];
    int s = 0;
    if (nks > 0) {
        load(0, a0, b0);
        load(min(1, nks - 1), a1, b1);
    }
    for (; s + 2 < nks; s += 3) {
        load(s + 2, a2, b2);
        step(a0, b0);
        load(min(s + 3, nks - 1), a0, b0);
        step(a1, b1);
        load(min(s + 4, nks - 1), a1, b1);
        step(a2, b2);
    }
    if (s < nks) step(a0, b0);
    if (s + 1 < nks) step(a1, b1);
    // output element v of a lane: row 4 v + lane / 16, column lane % 16 of the 16 x 16 block
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = groups[(long)blockIdx.x * G + 16 * rt + 4 * v + (lane >> 4)];
            if (r < 0) continue;
            double* const Mr = M + (long)(r - row_base) * ld;
#pragma unroll
            for (int ct = 0; ct < CPW; ++ct) {
                const int f = col0 + 16 * ct;
                if (f < NP) Mr[f] = acc[rt][ct][v];
            }
        }
}

// Stage 1 on FP64 MFMA with the l-space weights formed in place (no staged
// weights in HBM): as sep_rowmom_mfma with RT = 2 and 4 waves, over chunks of
// 16 l3 values.  Between a chunk's K-steps the workgroup forms the next
// chunk's 16 x 32 weights (8 / NW per thread), permsix h^2 v v v (exact h^2
// by sep_weight from the f table, or the asymptotic form by lspace_weight),
// into the other half of a double-buffered LDS tile; one barrier per chunk.
// The arithmetic is the staged kernel's, so the weights are bit-identical to
// the staged ones.  B operands are loaded 3 K-steps ahead into 4 register
// sets.  l-space only (l3 step 2).
// swap = 0: the rows of a group share l1 (its constant leg) and differ in
// l2; swap = 1: they share l2 and differ in l1.  The weight is symmetric in
// (l1, l2), so the table windows below hold for either leg; the row factor
// and the point where permsix changes (l3 = l2) are taken with the legs in
// (l1, l2) order, so the weights are the staged kernel's bit for bit.
template <int CPW, int ASYM, int NW, bool SWAP>
__global__ __launch_bounds__(64 * NW, (NW == 1 && CPW == 4) ? 3 : 2) void sep_rowmom_lw(const int* __restrict__ groups,
                                                        const int4* __restrict__ ginfo,
                                                        const RowDesc* __restrict__ rows,
                                                        const double* __restrict__ f,
                                                        const double* __restrict__ v, int row_base, int ngroups,
                                                        const double* __restrict__ U, int ldu, int NP, int ld,
                                                        double* __restrict__ M)
{
    constexpr bool swap = SWAP;
    constexpr int G = SR_GM, RT = SR_GM / 16;
    constexpr int WPT = 8 / NW;   // weights of a chunk per thread (64 NW threads)
    static_assert(WPT * 64 * NW == 16 * G, "a chunk's weights split evenly over the threads");
    __shared__ double Wt[2][16][G];
    __shared__ int gj[G], gk0[G], gnk[G];
    // one wave, exact h^2: a chunk's table reads come from four 64-entry
    // windows in LDS (f at A1.., A2.., A3.., 1/((2a+1) f) at A4..), loaded by
    // one coalesced read each, double-buffered by chunk
    constexpr bool WIN = !ASYM && NW == 1;
    __shared__ double Wf[WIN ? 2 : 1][4][64];
    // runs of SR_XC consecutive groups (overlapping l3 windows, so shared pair
    // products in L2) are dealt to the 8 XCDs in turn: blocks b, b + 8, ...
    // share an XCD
    const int slot = blockIdx.x >> 3;
    const int gb = ((blockIdx.x & 7) + 8 * (slot / SR_XC)) * SR_XC + slot % SR_XC;
    if (gb >= ngroups) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    // the group's l3 range, the leg its rows share (gi: l1, or l2 when swap)
    // and the range [jlo, jhi] of their other leg, from its table entry: the
    // table windows and the first basis rows are read without waiting for the
    // row descriptors
    const int4 info = ginfo[gb];
    const int kmin = info.x, nkk = info.y;
    const int gi = info.z, jlo = info.w & 0xffff, jhi = info.w >> 16;
    if (tid < G) {
        const int r = groups[(long)gb * G + tid];
        // padding rows: no points, the first row's legs (keeps window indices in range)
        const RowDesc d = r >= 0 ? rows[r] : RowDesc{0, 0, 0, 0};
        gj[tid] = r >= 0 ? (swap ? d.i : d.j) : jlo;   // the row's other leg
        gk0[tid] = d.k0;
        gnk[tid] = d.nk;
    }
    const int nks = (nkk + 3) >> 2;           // K-steps
    const int nch = (nkk + 15) >> 4;          // chunks of 4 K-steps
    // 4 column tiles: lane l of wave w holds columns 16 CPW w + CPW (l % 16)
    // + ct of tile ct (consecutive doubles: 16-byte loads, half the load
    // instructions; 2.85 -> 2.68 ms over 59 columns).  Wider waves keep column
    // 16 CPW w + 16 ct + l % 16 (the consecutive layout took 15.2 -> 16.3 ms
    // at 8 tiles).  Each output column sums in l3 order either way.
    constexpr bool CONS = CPW == 4;
    const int col0 = 16 * CPW * w + (CONS ? CPW * (lane & 15) : (lane & 15));
    // K-step s reads U row kmin + 2 (4 s + lane / 16); s stays below nks, so
    // rows reach at most 6 past the group's last l3, into the table's zero
    // padding (SR_UPAD rows after lmax), and meet zero weights there
    const double* const Ubl = U + (long)(kmin + 2 * (lane >> 4)) * ldu + col0;
    const long ustep4 = 8l * ldu;

    auto loadb = [&](int s, double (&b)[CPW]) {
        const double* Ur = Ubl + ustep4 * s;
        if constexpr (CONS) {
#pragma unroll
            for (int ct = 0; ct < CPW; ct += 2) {
                const fd2 u2 = *reinterpret_cast<const fd2*>(Ur + ct);
                b[ct] = u2.x;
                b[ct + 1] = u2.y;
            }
        } else {
#pragma unroll
            for (int ct = 0; ct < CPW; ++ct) b[ct] = Ur[16 * ct];
        }
    };

    fd4 acc[RT][CPW];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int ct = 0; ct < CPW; ++ct) acc[rt][ct] = (fd4){0.0, 0.0, 0.0, 0.0};
    auto mfmas = [&](const double (&a)[RT], const double (&b)[CPW]) {
#pragma unroll
        for (int ct = 0; ct < CPW; ++ct)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
                acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
    };
    double bs[4][CPW];
    // the first three K-steps' basis rows and chunks 0 and 1's windows, in
    // flight while the row descriptors arrive
    auto wload = [&](int c, double (&wr)[4]) {
        const int L0 = kmin + 32 * c;
        const int A1 = (jlo + L0 - gi) >> 1, A2 = (gi + L0 - jhi) >> 1, A3 = (gi + jlo - L0 - 30) >> 1,
                  A4 = (gi + jlo + L0) >> 1;
        wr[0] = f[2 * (A1 + lane)];
        wr[1] = f[2 * (A2 + lane)];
        wr[2] = f[2 * (A3 + lane)];
        wr[3] = f[2 * (A4 + lane) + 1];
    };
    auto wstore = [&](int b, const double (&wr)[4]) {
#pragma unroll
        for (int q = 0; q < 4; ++q) Wf[b][q][lane] = wr[q];
    };
    double wr0[WIN ? 4 : 1], wr1[WIN ? 4 : 1];
    if (nks > 0) {
#pragma unroll
        for (int x = 0; x < 3; ++x) loadb(min(x, nks - 1), bs[x]);
        if constexpr (WIN) {
            wload(0, wr0);
            wload(min(1, nch - 1), wr1);
        }
    }
    __syncthreads();   // row info
    // weights of chunk c into Wt[buf]: thread tid forms the elements
    // tid + 64 NW x (x < 8 / NW), element e = (kk = e / 32, g = e % 32),
    // l3 = kmin + 2 (16 c + kk);
    // all are of row g = tid % 32, whose constants stay in registers.  The
    // table reads of a group of two weights (gather) are issued one or more
    // K-steps of MFMAs before their use (weigh); exact weights by
    // sep_weight, as the staged ones.
    const int g = tid & (G - 1), kk0 = tid / G;   // element x: l3 index kk0 + 2 NW x of the chunk
    const int rj = gj[g], rk0 = gk0[g], rnk = gnk[g];
    // window geometry: the rows' j lie in [jlo, jhi] (one parity, span <= 96:
    // checked on the host), so for l3 = L0 + 2 t (t < 16) the f indices are
    // A + dj + t, A + dh + t, A + dj + 15 - t and the 1/((2a+1) f) index A +
    // dj + t, all below 64
    const int dj = (rj - jlo) >> 1, dh = (jhi - rj) >> 1;
    // the legs in (l1, l2) order, and the l3 at which permsix changes (l3 = l2)
    const int li = swap ? rj : gi, lj = swap ? gi : rj;
    // row factors of the weight at l3 = l2 and at l3 > l2 (l3 >= l2 >= l1 on the row)
    const double re = sep_row_factor(li, lj, permsix_d(li, lj, lj), v),
                 rn = sep_row_factor(li, lj, permsix_d(li, lj, lj + 1), v);
    constexpr int XS = 4 * NW;   // l3 (and table index) step between a thread's elements
    struct Tab {
        double f1, f2, f3, rg;
        int k;
        bool in;
    };
    // a thread's weights of a chunk are formed in groups of two (elements 2 gr
    // and 2 gr + 1), one group's table reads in flight at a time.  With one
    // wave (WIN) the reads come from the chunk's LDS windows.  With more
    // waves, every table index is affine in the element along the thread's
    // row (g = (i + j + l3) / 2 moves by 2 NW): the reads go through
    // per-chunk base pointers with constant offsets.  Either way the tables
    // carry zero guards (f: ftab_guard doubles before index 0; f and v:
    // TAB_TAIL after the end), so elements outside the row read in bounds and
    // are dropped by their flag.
    constexpr int NG = WPT / 2;
    auto gather = [&](int c, int gr, Tab (&t)[2]) {
        const int l30 = kmin + 32 * c + 2 * kk0;   // the thread's l3 at element 0
        const int u0 = ((kmin - rk0) >> 1) + 16 * c + kk0;   // its index on the row
        if (ASYM) {
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int x = 2 * gr + y;
                t[y].k = l30 + XS * x;
                t[y].in = (unsigned)(u0 + 2 * NW * x) < (unsigned)rnk;
            }
        } else if constexpr (WIN) {
            const double* const W = &Wf[c & 1][0][0];
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int x = 2 * gr + y;
                const int tt = kk0 + 2 * x;   // l3 = kmin + 32 c + 2 tt
                Tab& tx = t[y];
                tx.f1 = W[dj + tt];
                tx.f2 = W[64 + dh + tt];
                tx.f3 = W[128 + dj + 15 - tt];
                tx.rg = W[192 + dj + tt];
                tx.k = l30 + XS * x;
                tx.in = (unsigned)(u0 + 2 * NW * x) < (unsigned)rnk;
            }
        } else {
            const int g0 = (gi + rj + l30) >> 1;
            const double* const q1 = f + 2 * (g0 - gi);
            const double* const q2 = f + 2 * (g0 - rj);
            const double* const q3 = f + 2 * (g0 - l30);
            const double* const q4 = f + 2 * g0 + 1;
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int x = 2 * gr + y;
                Tab& tx = t[y];
                tx.f1 = q1[XS * x];
                tx.f2 = q2[XS * x];
                tx.f3 = q3[-XS * x];
                tx.rg = q4[XS * x];
                tx.k = l30 + XS * x;
                tx.in = (unsigned)(u0 + 2 * NW * x) < (unsigned)rnk;
            }
        }
    };
    auto weigh = [&](const Tab (&t)[2], int gr, int buf) {
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int x = 2 * gr + y;
            const Tab& tx = t[y];
            double wv;
            if (ASYM) {
                wv = lspace_weight(tx.in ? li : 0, tx.in ? lj : 0, tx.in ? tx.k : 0, 1, f, v);
            } else {
                wv = sep_weight(tx.k == lj ? re : rn, tx.f1, tx.f2, tx.f3, tx.rg);
            }
            Wt[buf][kk0 + 2 * NW * x][g] = tx.in ? wv : 0.0;
        }
    };
    // chunks after the first of a group sharing l2 (exact h^2): every row
    // starts at l3 = kmin (its l1 parity is the group's), so no element lies
    // below its row; above it the tables' zero guards make the weight +0
    // exactly (f(g - l3) with g < l3, or v past lmax), as the range flag
    // would; and l3 = l2 only at the first element of chunk 0, so the row
    // factor is rn.  Same products, same bits as weigh.
    auto weigh_tail = [&](const Tab (&t)[2], int gr, int buf) {
        if constexpr (SWAP && !ASYM) {
#pragma unroll
            for (int y = 0; y < 2; ++y) {
                const int x = 2 * gr + y;
                const Tab& tx = t[y];
                Wt[buf][kk0 + 2 * NW * x][g] = sep_weight(rn, tx.f1, tx.f2, tx.f3, tx.rg);
            }
        } else {
            weigh(t, gr, buf);
        }
    };
    if constexpr (WIN) {
        if (nks > 0) {
            wstore(0, wr0);
            wstore(1, wr1);
        }
        __syncthreads();
    }
    if (nks > 0) {
        Tab t0[2];
#pragma unroll
        for (int gr = 0; gr < NG; ++gr) {
            gather(0, gr, t0);
            weigh(t0, gr, 0);
        }
    }
    __syncthreads();
    // all chunks but the last: 4 K-steps each; the next chunk's table reads
    // are issued first and its weights formed after the third step's MFMAs
    int c = 0;
    for (; c + 1 < nch; ++c) {
        const int buf = c & 1;
        double a[4][RT];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int rt = 0; rt < RT; ++rt) a[ks][rt] = Wt[buf][4 * ks + (lane >> 4)][16 * rt + (lane & 15)];
        // windows of chunk c + 2, stored into this chunk's buffer at its end
        double wr[WIN ? 4 : 1];
        if constexpr (WIN) wload(min(c + 2, nch - 1), wr);
        Tab t[2];
        gather(c + 1, 0, t);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            loadb(min(4 * c + ks + 3, nks - 1), bs[(ks + 3) & 3]);
            mfmas(a[ks], bs[ks]);
            // group gr's arithmetic after the MFMAs of step 2 (one group),
            // steps 0 and 2 (two) or steps 0..3 (four), the next group's
            // reads right behind it
            const int gr = NG == 1 ? (ks == 2 ? 0 : -1) : NG == 2 ? (ks == 0 ? 0 : ks == 2 ? 1 : -1) : ks;
            if (gr >= 0) {
                __builtin_amdgcn_sched_barrier(0);
                weigh_tail(t, gr, buf ^ 1);
                if (gr + 1 < NG) gather(c + 1, gr + 1, t);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        if constexpr (WIN) wstore(buf, wr);
        __syncthreads();
    }
    // the last chunk: its K-steps below nks (steps 4c .. 4c + 2 are loaded;
    // step 4c + 3 is loaded here)
    if (c < nch) {
        const int buf = c & 1;
        loadb(min(4 * c + 3, nks - 1), bs[3]);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            if (4 * c + ks < nks) {
                double a[RT];
#pragma unroll
                for (int rt = 0; rt < RT; ++rt) a[rt] = Wt[buf][4 * ks + (lane >> 4)][16 * rt + (lane & 15)];
                mfmas(a, bs[ks]);
            }
        }
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
        for (int vv = 0; vv < 4; ++vv) {
            const int r = groups[(long)gb * G + 16 * rt + 4 * vv + (lane >> 4)];
            if (r < 0) continue;
            double* const Mr = M + (long)(r - row_base) * ld;
#pragma unroll
            for (int ct = 0; ct < CPW; ++ct) {
                const int fcol = col0 + (CONS ? ct : 16 * ct);
                if (fcol < NP) Mr[fcol] = acc[rt][ct][vv];
            }
        }
}

// Stages 2 and 3:  C_task[m][n] (+)= sum_{kk < K} U[row(kk)][m] * B[boff + kk][n]
// for m < Md, n < Nd, over a list of 160 x 96 tiles (task, m0, n0); row(kk) =
// aoff + kk when the A rows are contiguous (CONTIG), else aidx[aoff + kk].
// Both operands are K-major with even leading dimensions (16-byte loads; the
// element past an odd Md or Nd is padding and only feeds outputs that are not
// stored).  16 K rows per stage, double-buffered in LDS; FP64 MFMA 16x16x4
// with 4 waves of 80 x 48.  The tile list holds groups of GS consecutive
// tiles that share one B panel (the m tiles of one (task, n0)).  Blocks b,
// b+8, ... share an XCD (and its L2): the groups are dealt to the 8 XCDs in
// turn, each group's tiles to consecutive blocks of one XCD, so the tiles of
// a group meet in the same L2 and the XCDs get alike shares of the work.
// STAGE (2: l2, 3: l1, 4: expansion) only names the launches apart in
// profiles.  FM x FN: 16 x 16 blocks per wave (5 x 3: 160 x 96 tiles; 2 x 2:
// 64 x 64 tiles for narrow bases).  PAIRS: C is symmetric (Md = Nd) and only
// its upper triangle is stored, at C + coff + pair_index(m, n, Md).
template <bool CONTIG, int STAGE, int FM = SG_FM, int FN = SG_FN, bool PAIRS = false>
__global__ __launch_bounds__(256, 2) void sep_gemm(const double* __restrict__ U, int ldu,
                                                   const int* __restrict__ aidx, const double* __restrict__ B,
                                                   long ldb, double* __restrict__ C, long ldc, int Md, int Nd,
                                                   const SepTask* __restrict__ tasks, const int4* __restrict__ tiles,
                                                   int ntiles, int GS, int accumulate)
{
    constexpr int TM = 32 * FM, TN = 32 * FN;
    constexpr int LDA = TM + 16, LDB = TN + 16;   // LDS row strides in doubles (32 banks apart)
    __shared__ double As[2][SG_KC][LDA];
    __shared__ double Bs[2][SG_KC][LDB];
    const int slot = blockIdx.x >> 3;
    const int t = ((blockIdx.x & 7) + 8 * (slot / GS)) * GS + slot % GS;
    if (t >= ntiles) return;
    const int4 tl = tiles[t];
    const SepTask tk = tasks[tl.x];
    const int m0 = tl.y, n0 = tl.z, K = tk.K;
    const int* ai = aidx + tk.aoff;
    const double* Bb = B + tk.boff * ldb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;

    fd4 acc[FM][FN];
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int c = 0; c < FN; ++c) acc[a][c] = (fd4){0.0, 0.0, 0.0, 0.0};

    constexpr int HA = TM / 2, HB = TN / 2;   // double2 per tile row
    constexpr int PA = SG_KC * HA / 256;             // double2 of A per thread
    constexpr int PB = SG_KC * HB / 256;             // double2 of B per thread
    static_assert(PA * 256 == SG_KC * HA && PB * 256 == SG_KC * HB, "tile loads split evenly over 256 threads");
    const fd2 z2 = {0.0, 0.0};
    fd2 ra[PA], rb[PB];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            const int e = tid + 256 * u;
            const int kk = e / HA, mm = 2 * (e - HA * kk);
            const int kg = k0 + kk;
            const bool kin = kg < K;
            const long arow = CONTIG ? (long)(tk.aoff + kg) : (long)(kin ? ai[kg] : 0);
            ra[u] = (kin && m0 + mm < Md) ? *reinterpret_cast<const fd2*>(U + arow * ldu + m0 + mm) : z2;
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int e = tid + 256 * u;
            const int kk = e / HB, nn = 2 * (e - HB * kk);
            const int kg = k0 + kk;
            rb[u] = (kg < K && n0 + nn < Nd) ? *reinterpret_cast<const fd2*>(Bb + (long)kg * ldb + n0 + nn) : z2;
        }
    };
    auto sstore = [&](int b) {
#pragma unroll
        for (int u = 0; u < PA; ++u) {
            const int e = tid + 256 * u;
            const int kk = e / HA, mm = 2 * (e - HA * kk);
            *reinterpret_cast<fd2*>(&As[b][kk][mm]) = ra[u];
        }
#pragma unroll
        for (int u = 0; u < PB; ++u) {
            const int e = tid + 256 * u;
            const int kk = e / HB, nn = 2 * (e - HB * kk);
            *reinterpret_cast<fd2*>(&Bs[b][kk][nn]) = rb[u];
        }
    };

    const int nch = (K + SG_KC - 1) / SG_KC;
    if (nch > 0) {
        gload(0);
        sstore(0);
        __syncthreads();
        for (int ch = 0; ch < nch; ++ch) {
            const int b = ch & 1;
            if (ch + 1 < nch) gload((ch + 1) * SG_KC);
#pragma unroll
            for (int ks = 0; ks < SG_KC / 4; ++ks) {
                const int kk = ks * 4 + (lane >> 4);
                double av[FM], bv[FN];
#pragma unroll
                for (int a = 0; a < FM; ++a) av[a] = As[b][kk][wm * (TM / 2) + a * 16 + (lane & 15)];
#pragma unroll
                for (int c = 0; c < FN; ++c) bv[c] = Bs[b][kk][wn * (TN / 2) + c * 16 + (lane & 15)];
#pragma unroll
                for (int a = 0; a < FM; ++a)
#pragma unroll
                    for (int c = 0; c < FN; ++c)
                        acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
            }
            if (ch + 1 < nch) sstore(b ^ 1);
            __syncthreads();
        }
    }
    // output element v of a lane: row 4v + lane/16, column lane%16 of the 16 x 16 block
    double* const Cb = C + tk.coff;
#pragma unroll
    for (int a = 0; a < FM; ++a)
#pragma unroll
        for (int c = 0; c < FN; ++c) {
            const int n = n0 + wn * (TN / 2) + c * 16 + (lane & 15);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm * (TM / 2) + a * 16 + 4 * v + (lane >> 4);
                if (m < Md && n < Nd && (!PAIRS || m <= n)) {
                    double* p = Cb + (PAIRS ? (long)pair_index(m, n, Md) : (long)m * ldc + n);
                    *p = accumulate ? *p + acc[a][c][v] : acc[a][c][v];
                }
            }
        }
}

// Ys[i][{e,f}] = Y[i][e][f] + Y[i][f][e] for e <= f ({e,f} = pair_index(e, f,
// NP)): one workgroup per (32 x 32 block pair E <= F, run i); both blocks are
// read row by row (coalesced), the second transposed through LDS.  The
// padding element of each Ys row (index NP(NP+1)/2 when that is odd) is
// zeroed.
__global__ __launch_bounds__(256) void sep_ysym(const double* __restrict__ Y, long ldyi, int ld, int NP, int nb,
                                                double* __restrict__ Ys, long ldys)
{
    __shared__ double A[32][33], Bt[32][33];
    int bE = 0, rem = blockIdx.x;   // block pair (bE <= bF) in row-major upper-triangle order
    while (rem >= nb - bE) {
        rem -= nb - bE;
        ++bE;
    }
    const int bF = bE + rem;
    const int e0 = 32 * bE, f0 = 32 * bF;
    const double* const Yi = Y + (long)blockIdx.y * ldyi;
    double* const Yo = Ys + (long)blockIdx.y * ldys;
    const int tid = threadIdx.x, c = tid & 31;
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
        const int r = (tid >> 5) + 8 * r4;
        A[r][c] = (e0 + r < NP && f0 + c < NP) ? Yi[(long)(e0 + r) * ld + f0 + c] : 0.0;
        Bt[r][c] = (f0 + r < NP && e0 + c < NP) ? Yi[(long)(f0 + r) * ld + e0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r4 = 0; r4 < 4; ++r4) {
        const int r = (tid >> 5) + 8 * r4;
        const int e = e0 + r, f = f0 + c;
        if (e < NP && f < NP && e <= f) Yo[pair_index(e, f, NP)] = A[r][c] + Bt[c][r];
    }
    const long npp = (long)NP * (NP + 1) / 2;
    if (blockIdx.x == 0 && tid == 0 && ldys > npp) Yo[npp] = 0.0;
}

// Gs[off[{x,y}] + z - y] = G2[x][{y,z}] + G2[y][{x,z}] + G2[z][{x,y}] for
// x <= y <= z: one workgroup per pair {x, y}, threads over z.
__global__ __launch_bounds__(256) void sep_gsym(const double* __restrict__ G2, long ldg, int NP,
                                                const long* __restrict__ off, double* __restrict__ Gs)
{
    const int p = blockIdx.x;
    int lo = 0, hi = NP - 1;   // x = the largest row whose first pair index is <= p
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pair_index(mid, mid, NP) <= p) lo = mid;
        else hi = mid - 1;
    }
    const int x = lo, y = x + (p - pair_index(x, x, NP));
    const double* const Gx = G2 + (long)x * ldg + pair_index(y, y, NP);   // [{y, z}] at z - y
    const double* const Gy = G2 + (long)y * ldg + pair_index(x, x, NP) - x;   // [{x, z}] at z
    const long pxy = pair_index(x, y, NP);
    double* const out = Gs + off[p];
    for (int z = y + threadIdx.x; z < NP; z += blockDim.x)
        out[z - y] = Gx[z - y] + Gy[z] + G2[(long)z * ldg + pxy];
}

// ---- compressed pair basis ----
//
// The pair products U_f(l) = q_x(l) q_y(l) of a basis of R functions often
// span far fewer than R(R+1)/2 directions in l (products of polynomials of
// degree < R are polynomials of degree < 2R - 1).  Pivoted modified
// Gram-Schmidt over the columns finds U = P Ct + E with P orthonormal (r
// columns) and |E_f| below a relative tolerance; the three contractions then
// run over the r columns of P, giving the symmetrised tensor H over P, and the
// expansion Gs(x, y, z) = sum_{a,b,g} Ct[a][x] Ct[b][y] Ct[g][z] H[a][b][g]
// returns to pairs.

// sum over the 64 lanes of a wave (butterfly, fixed order)
__device__ __forceinline__ double wave_sum(double x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

// Uc[f][l] = q_x(l) q_y(l), f = pair (x, y): the pair products as columns
// (leading dimension nlp; rows l >= nl zero).
__global__ void sep_ucols(const double* __restrict__ q, int ldq, int nl, int nlp, const int2* __restrict__ pxy,
                          int NP, double* __restrict__ Uc)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)NP * nlp) return;
    const int f = (int)(e / nlp), l = (int)(e % nlp);
    const int2 xy = pxy[f];
    Uc[e] = l < nl ? q[(long)l * ldq + xy.x] * q[(long)l * ldq + xy.y] : 0.0;
}

// nrm[f] = |Uc[f]|^2, one wave per column
__global__ __launch_bounds__(256) void sep_mgs_norms(const double* __restrict__ Uc, int nlp, int NP,
                                                     double* __restrict__ nrm)
{
    const int lane = threadIdx.x & 63;
    const int f = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (f >= NP) return;
    const double* const u = Uc + (long)f * nlp;
    double s = 0.0;
    for (int l = lane; l < nlp; l += 64) s += u[l] * u[l];
    s = wave_sum(s);
    if (lane == 0) nrm[f] = s;
}

// Step a of pivoted modified Gram-Schmidt on the residual columns Uc (updated
// in place).  nin[f] is the squared residual norm of column f after step a - 1
// (-1 once f has been a pivot); the pivot p is the largest (the smallest f on
// ties), and its squared norm is recorded in piv[a].  The factorisation stops
// (rank[0] = a) when the pivot's squared norm is at most tol2 times the first
// pivot's (kept in n0[0]), and gives up (rank[0] = -2) at a = amax.  Otherwise, with v = Uc[p] / |Uc[p]|: Pc[a] = v;
// for every column that has not been a pivot Ct[a][f] = v . Uc[f], Uc[f] -=
// Ct[a][f] v and nout[f] = |Uc[f]|^2; Ct[a][p] = |Uc[p]| and Ct[a][f] = 0 for
// earlier pivots.  Every workgroup takes the same decisions from the same nin
// and the pivot column is only read, so the in-place update is race-free.
// Launched for a = 0, 1, ... in stream order; after rank[0] is set the later
// launches return at once.  One wave per column.
__global__ __launch_bounds__(256) void sep_mgs_step(double* __restrict__ Uc, int nlp, int NP,
                                                    const double* __restrict__ nin, double* __restrict__ nout,
                                                    int a, int amax, double tol2, double* __restrict__ n0,
                                                    int* __restrict__ rank, double* __restrict__ piv,
                                                    double* __restrict__ Pc, double* __restrict__ Ct, int ldct)
{
    __shared__ double sv[256];
    __shared__ int si[256];
    if (*rank != -1) return;
    const int tid = threadIdx.x;
    double best = -2.0;
    int bi = NP;
    for (int f = tid; f < NP; f += 256) {
        const double x = nin[f];
        if (x > best) {
            best = x;
            bi = f;
        }
    }
    sv[tid] = best;
    si[tid] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) {
            const double o = sv[tid + s];
            const int oi = si[tid + s];
            if (o > sv[tid] || (o == sv[tid] && oi < si[tid])) {
                sv[tid] = o;
                si[tid] = oi;
            }
        }
        __syncthreads();
    }
    const double n2 = sv[0];
    const int p = si[0];
    const double ref = a == 0 ? n2 : n0[0];
    if (blockIdx.x == 0 && tid == 0) piv[a] = n2;
    if (!(n2 > tol2 * ref)) {
        if (blockIdx.x == 0 && tid == 0) *rank = a;
        return;
    }
    if (a >= amax) {
        if (blockIdx.x == 0 && tid == 0) *rank = -2;
        return;
    }
    if (a == 0 && blockIdx.x == 0 && tid == 0) n0[0] = n2;
    const double nrm = sqrt(n2), inv = 1.0 / nrm;
    const double* const pv = Uc + (long)p * nlp;
    if (blockIdx.x == 0)
        for (int l = tid; l < nlp; l += 256) Pc[(long)a * nlp + l] = pv[l] * inv;
    const int lane = tid & 63;
    for (int f = blockIdx.x * 4 + (tid >> 6); f < NP; f += gridDim.x * 4) {
        if (f == p || nin[f] < 0.0) {
            if (lane == 0) {
                Ct[(long)a * ldct + f] = f == p ? nrm : 0.0;
                nout[f] = -1.0;
            }
            continue;
        }
        double* const u = Uc + (long)f * nlp;
        double d = 0.0;
        for (int l = lane; l < nlp; l += 64) d += (pv[l] * inv) * u[l];
        d = wave_sum(d);
        double s = 0.0;
        for (int l = lane; l < nlp; l += 64) {
            const double x = u[l] - d * (pv[l] * inv);
            u[l] = x;
            s += x * x;
        }
        s = wave_sum(s);
        if (lane == 0) {
            Ct[(long)a * ldct + f] = d;
            nout[f] = s;
        }
    }
}

// P[l][a] = Pc[a][l] for a < r, 0 for r <= a < ldu: the compressed basis in
// the row layout of the pair products
// The l3 stage's basis table for the exact l-space weight: U3[l] = (2l + 1)
// v_l U[l] (the weight's l3 factor, sep_l3_factor), zero past the last l.
__global__ void sep_u3tab(const double* __restrict__ U, int ldu, int nl, const double* __restrict__ v,
                          double* __restrict__ U3)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(nl + SR_UPAD) * ldu) return;
    const int l = (int)(e / ldu);
    U3[e] = l < nl ? sep_l3_factor(l, v) * U[e] : 0.0;
}

__global__ void sep_ptab(const double* __restrict__ Pc, int nlp, int nl, int r, int ldu, double* __restrict__ P)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)(nl + SR_UPAD) * ldu) return;
    const int l = (int)(e / ldu), a = (int)(e % ldu);
    P[e] = (a < r && l < nl) ? Pc[(long)a * nlp + l] : 0.0;
}

// H[g][a r + b] = Hs(sorted (g, a, b)) with Hs(x, y, z) = G2[x][{y,z}] +
// G2[y][{x,z}] + G2[z][{x,y}] (x <= y <= z, pairs over r), each G2 entry the
// sum of the ns partial sums G2 + s lds in order of s: every entry of the
// fully symmetrised tensor over the compressed basis, equal for all orders of
// (g, a, b); the padding entry of each row (r^2 odd) is zeroed.
__global__ void sep_hfull(const double* __restrict__ G2, long ldg, long lds, int ns, int r, double* __restrict__ H,
                          long ldh)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)r * ldh) return;
    const int g = (int)(e / ldh);
    const long ab = e % ldh;
    if (ab >= (long)r * r) {
        H[e] = 0.0;
        return;
    }
    const int a = (int)(ab / r), b = (int)(ab % r);
    const int x = min(g, min(a, b)), z = max(g, max(a, b)), y = g + a + b - x - z;
    auto g2 = [&](int u, int p) {
        double t = 0.0;
        for (int sp = 0; sp < ns; ++sp) t += G2[sp * lds + (long)u * ldg + p];
        return t;
    };
    H[e] = g2(x, pair_index(y, z, r)) + g2(y, pair_index(x, z, r)) + g2(z, pair_index(x, y, r));
}

// Gamma[m][n] = Gamma[n][m] = (1/36) sum over the 6 matchings sigma of the
// slots of m and n of Gs at the sorted pair triple.
__global__ void sep_assemble(const double* __restrict__ Gs, const long* __restrict__ off, int NP, int R,
                             const int* __restrict__ modes, int n, double* __restrict__ gamma)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int m = (int)(e / n), nn = (int)(e % n);
    if (m > nn) return;
    const int pm = modes[m], pn = modes[nn];
    const int a[3] = {pm & 255, (pm >> 8) & 255, (pm >> 16) & 255};
    const int b[3] = {pn & 255, (pn >> 8) & 255, (pn >> 16) & 255};
    int pid[3][3];
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) pid[x][y] = pair_index(min(a[x], b[y]), max(a[x], b[y]), R);
    constexpr int P6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s < 6; ++s) {
        int u = pid[0][P6[s][0]], v = pid[1][P6[s][1]], w = pid[2][P6[s][2]];
        const int x = min(u, min(v, w)), z = max(u, max(v, w)), y = u + v + w - x - z;
        acc += Gs[off[pair_index(x, y, NP)] + (z - y)];
    }
    acc *= (1.0 / 36.0);
    gamma[(long)m * n + nn] = acc;
    gamma[(long)nn * n + m] = acc;
}

// Gamma from the fully symmetrised tensor G[x][{y,z}] (the expansion's
// output, computed where x <= y), over 32 x 32 tiles (bm <= bn) of Gamma:
// for each matching sigma the sorted triple x <= y <= z of its pairs is read.
// Entries m <= nn are computed; the tile and (off the diagonal) its mirror
// are written from LDS along rows, so Gamma is exactly symmetric.  The reads
// of G are gathers (about 0.6 distinct 128-byte lines per entry of a tile at
// 4960 modes), so neighbouring tiles are kept on one XCD (S x S blocks, see
// below) to share its L2.
template <int S>
__global__ __launch_bounds__(256) void sep_assemble_sym(const double* __restrict__ G, long ldg, int R,
                                                        const int* __restrict__ modes, int n, int nt,
                                                        double* __restrict__ gamma)
{
    __shared__ double T[32][33];
    // tile (bm <= bn) of this workgroup: S x S blocks of tiles, blocks in
    // row-major upper-triangle order, block k on XCD k % 8 with its S^2 tiles
    // in consecutive dispatch slots there (workgroups b, b + 8, ... share an
    // XCD), so the tiles that read largely the same lines of G run together
    // and share that XCD's L2
    const int nb = (nt + S - 1) / S;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    int rem = (slot / (S * S)) * 8 + xcd;
    const int sub = slot % (S * S);
    if (rem >= nb * (nb + 1) / 2) return;
    int sm = 0;
    while (rem >= nb - sm) {
        rem -= nb - sm;
        ++sm;
    }
    const int bm = S * sm + sub / S, bn = S * (sm + rem) + sub % S;
    if (bm > bn || bn >= nt) return;
    const int c = threadIdx.x & 31, r0 = threadIdx.x >> 5;
    const int nn = 32 * bn + c;
    const int pn = nn < n ? modes[nn] : 0;
    const int b[3] = {pn & 255, (pn >> 8) & 255, (pn >> 16) & 255};
    constexpr int P6[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 8 * i, m = 32 * bm + r;
        double acc = 0.0;
        if (m < n && nn < n && m <= nn) {
            const int pm = modes[m];
            const int a[3] = {pm & 255, (pm >> 8) & 255, (pm >> 16) & 255};
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                // slot x of m meets slot P6[s][x] of n
                const int u = pair_index(min(a[0], b[P6[s][0]]), max(a[0], b[P6[s][0]]), R);
                const int v = pair_index(min(a[1], b[P6[s][1]]), max(a[1], b[P6[s][1]]), R);
                const int w = pair_index(min(a[2], b[P6[s][2]]), max(a[2], b[P6[s][2]]), R);
                const int x = min(u, min(v, w)), z = max(u, max(v, w)), y = u + v + w - x - z;
                acc += G[(long)x * ldg + pair_index(y, z, R * (R + 1) / 2)];
            }
            acc *= (1.0 / 36.0);
        }
        T[r][c] = acc;
    }
    __syncthreads();
    // Gamma is written once and not read here: streaming stores keep the
    // caches for G
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 8 * i, m = 32 * bm + r;
        if (m < n && nn < n) __builtin_nontemporal_store(m <= nn ? T[r][c] : T[c][r], &gamma[(long)m * n + nn]);
    }
    if (bm != bn) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + 8 * i, row = 32 * bn + r, col = 32 * bm + c;
            if (row < n && col < n) __builtin_nontemporal_store(T[c][r], &gamma[(long)row * n + col]);
        }
    }
}

// A polarised Gamma from the mixed GL expansion K[x][{y, z}] (x a pair of
// the field on one leg, {y, z} two pairs of the field on the other two legs,
// in pair order), over 32 x 32 tiles (bm <= bn) placed as in
// sep_assemble_sym, with p_xy = pair(a_x, b_y) of T and u_xy of E:
//   KIND 1 (TTE, legs T T E): K[u22][{p00, p11}] + K[u22][{p10, p01}]
//   KIND 2 (TEE, legs T E E): K[p00][{u11, u22}] + K[p00][{u21, u12}]
// (the two matchings of orthogonalisation_late.c:302-316; the 1/(16 pi) is
// in the core).  Both sums are symmetric in m and n.
template <int KIND, int S>
__global__ __launch_bounds__(256) void sep_assemble_mix(const double* __restrict__ K, long ldk, int RT, int RE,
                                                        const int* __restrict__ modes, int n, int nt,
                                                        double* __restrict__ gamma)
{
    __shared__ double T[32][33];
    const int nb = (nt + S - 1) / S;
    const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
    int rem = (slot / (S * S)) * 8 + xcd;
    const int sub = slot % (S * S);
    if (rem >= nb * (nb + 1) / 2) return;
    int sm = 0;
    while (rem >= nb - sm) {
        rem -= nb - sm;
        ++sm;
    }
    const int bm = S * sm + sub / S, bn = S * (sm + rem) + sub % S;
    if (bm > bn || bn >= nt) return;
    const int c = threadIdx.x & 31, r0 = threadIdx.x >> 5;
    const int nn = 32 * bn + c;
    const int pn = nn < n ? modes[nn] : 0;
    const int b0 = pn & 255, b1 = (pn >> 8) & 255, b2 = (pn >> 16) & 255;
    auto pr = [](int x, int y, int R) { return pair_index(min(x, y), max(x, y), R); };
    const int NPP = KIND == 1 ? RT * (RT + 1) / 2 : RE * (RE + 1) / 2;   // pairs of the paired field
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 8 * i, m = 32 * bm + r;
        double acc = 0.0;
        if (m < n && nn < n && m <= nn) {
            const int pm = modes[m];
            const int a0 = pm & 255, a1 = (pm >> 8) & 255, a2 = (pm >> 16) & 255;
            int x, y1, z1, y2, z2;
            if (KIND == 1) {
                x = pr(a2, b2, RE);
                y1 = pr(a0, b0, RT);
                z1 = pr(a1, b1, RT);
                y2 = pr(a1, b0, RT);
                z2 = pr(a0, b1, RT);
            } else {
                x = pr(a0, b0, RT);
                y1 = pr(a1, b1, RE);
                z1 = pr(a2, b2, RE);
                y2 = pr(a2, b1, RE);
                z2 = pr(a1, b2, RE);
            }
            const double* const Kx = K + (long)x * ldk;
            acc = Kx[pr(y1, z1, NPP)] + Kx[pr(y2, z2, NPP)];
        }
        T[r][c] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int r = r0 + 8 * i, m = 32 * bm + r;
        if (m < n && nn < n) __builtin_nontemporal_store(m <= nn ? T[r][c] : T[c][r], &gamma[(long)m * n + nn]);
    }
    if (bm != bn) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = r0 + 8 * i, row = 32 * bn + r, col = 32 * bm + c;
            if (row < n && col < n) __builtin_nontemporal_store(T[c][r], &gamma[(long)row * n + col]);
        }
    }
}

bool gram_separable_supported(const PointSet& L, int precision)
{
    // (group table entries keep two legs in 16 bits each)
    return precision == MCMB_FP64 && L.R >= 1 && L.R <= 255 && L.R * (L.R + 1) / 2 <= SR_NPMAX && L.lmax < 32768;
}

// Row groups sharing l2 formed on the device, one wave per (l2, parity)
// class cls = 2 (j - ia) + par: the rows (i, j) of the block's l1 values i of
// parity par, i ascending, whose runs reach l2 = j (j <= jl_of[i - ia]), row
// index r0_of[i - ia] + j - i, in groups of G from group gstart[cls] on, the
// last one padded with -1 (the host walk's order and padding).  Then per
// group: its l3 range (ginfo), its staged-weight size (into wofs, scanned
// later) and the span of l1 over its rows (the largest into *jspan).
template <int G>
__global__ __launch_bounds__(64) void sep_groups_l2(const int* __restrict__ r0_of, const int* __restrict__ jl_of,
                                                    const int* __restrict__ gstart, int ia, int ib,
                                                    const RowDesc* __restrict__ rows, int kstep,
                                                    int* __restrict__ groups, int4* __restrict__ ginfo,
                                                    long* __restrict__ wofs, int* __restrict__ jspan)
{
    static_assert(G == 32, "two groups per wave pass");
    const int cls = blockIdx.x, j = ia + (cls >> 1), par = cls & 1;
    const int g0 = gstart[cls], g1 = gstart[cls + 1];
    if (g1 == g0) return;   // the whole workgroup returns
    const int lane = threadIdx.x;
    int* const out = groups + (long)g0 * G;
    const int ilast = min(j, ib);
    int pos = 0;
    for (int i0 = ia + ((ia & 1) != par); i0 <= ilast; i0 += 128) {
        const int i = i0 + 2 * lane;
        const bool ok = i <= ilast && j <= jl_of[i - ia];
        const unsigned long long m = __ballot(ok);
        if (ok) out[pos + __popcll(m & ((1ull << lane) - 1))] = r0_of[i - ia] + (j - i);
        pos += __popcll(m);
    }
    for (long e = pos + lane; e < (long)(g1 - g0) * G; e += 64) out[e] = -1;
    __syncthreads();
    const int h = lane / G, t = lane % G;
    for (int gp = g0; gp < g1; gp += 2) {
        const int g = gp + h;
        int kmin = INT_MAX, kmax = INT_MIN, di = 0, d0i = 0;
        if (g < g1) {
            const int r = out[(long)(g - g0) * G + t];
            if (r >= 0) {
                const RowDesc d = rows[r], d0 = rows[out[(long)(g - g0) * G]];
                di = d.i - d0.i;
                d0i = d0.i;
                if (d.nk > 0) {
                    kmin = d.k0;
                    kmax = d.k0 + kstep * (d.nk - 1);
                }
            }
        }
#pragma unroll
        for (int o = G / 2; o >= 1; o >>= 1) {
            kmin = min(kmin, __shfl_xor(kmin, o));
            kmax = max(kmax, __shfl_xor(kmax, o));
            di = max(di, __shfl_xor(di, o));
        }
        if (g < g1 && t == 0) {
            const int nkk = kmax >= kmin ? (kmax - kmin) / kstep + 1 : 0;
            // shared leg l2 = j; l1 from the first row's to the largest
            ginfo[g] = make_int4(nkk ? kmin : 0, nkk, j, d0i | ((d0i + di) << 16));
            wofs[g] = (long)G * 16 * ((nkk + 15) / 16);
            atomicMax(jspan, di);
        }
    }
}

// In-place exclusive prefix sum of the n group weight sizes in wofs (one
// workgroup: a contiguous range per thread, the range sums scanned in LDS);
// the sum of all into *total.
__global__ __launch_bounds__(1024) void sep_wofs_scan(long* __restrict__ wofs, long n, long* __restrict__ total)
{
    __shared__ long part[1024];
    const int t = threadIdx.x;
    const long per = (n + 1023) / 1024;
    const long a = min(n, t * per), b = min(n, a + per);
    long s = 0;
    for (long e = a; e < b; ++e) s += wofs[e];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const long x = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += x;
        __syncthreads();
    }
    long base = part[t] - s;
    for (long e = a; e < b; ++e) {
        const long x = wofs[e];
        wofs[e] = base;
        base += x;
    }
    if (t == 1023) *total = part[1023];
}

namespace {

struct Run {
    int r0, r1, i;
};

// One block's tables, one int array with each table 16-byte aligned.  Built
// on the host (buf holds all of it, host0 = 0), or with the row groups, their
// l3 ranges and weight offsets formed on the device (dev: buf holds only the
// small tables [host0, size); the device part needs, per l1 of the block, its
// first row and last l2, and the first group of each (l2, parity) class).
struct BlockTables {
    SepBlock b;
    std::unique_ptr<int[]> buf;
    size_t size = 0;    // ints of the block's tables
    size_t host0 = 0;   // first int that buf holds
    bool dev = false;
    int ia = 0, ib = 0, jtop = 0;
    std::vector<int> r0_of, jl_of, gstart;
};

template <class T> size_t append(std::vector<int>& buf, const std::vector<T>& v)
{
    static_assert(sizeof(T) % sizeof(int) == 0, "table entries are whole ints");
    while ((buf.size() * sizeof(int)) % 16) buf.push_back(0);   // 16-byte alignment
    const size_t off = buf.size();
    buf.resize(off + v.size() * sizeof(T) / sizeof(int));
    if (!v.empty()) memcpy(buf.data() + off, v.data(), v.size() * sizeof(T));
    return off;
}

// The l2 and l1 contraction tasks and tiles, and the l1 of each run.
struct SmallTables {
    std::vector<int> aidx3;
    std::vector<SepTask> task2, task3;
    std::vector<int4> tile2, tile3;
};

static void small_tables(const PointSet& L, const std::vector<Run>& runs, int b0, int NP, int ld, int tm, int tn,
                         int ksplit, SepBlock& T, SmallTables& S)
{
    S.aidx3.resize(T.nruns);
    S.task2.resize(T.nruns);
    bool c3 = true;
    for (int lt = 0; lt < T.nruns; ++lt) {
        const Run& R = runs[b0 + lt];
        S.task2[lt].boff = R.r0 - T.row0;
        S.task2[lt].coff = (long)lt * NP * ld;
        // contiguous A rows: the task's first row index replaces the list offset
        S.task2[lt].aoff = T.contig2 ? L.rows[R.r0].j : R.r0 - T.row0;
        S.task2[lt].K = R.r1 - R.r0;
        S.aidx3[lt] = R.i;
        if (R.i != runs[b0].i + lt) c3 = false;
        for (int n0 = 0; n0 < NP; n0 += tn)
            for (int m0 = 0; m0 < NP; m0 += tm) S.tile2.push_back(make_int4(lt, m0, n0, 0));
    }
    T.contig3 = c3;
    const long N3 = (long)NP * (NP + 1) / 2;   // stage 3 runs over the unordered pairs {e, f}
    const long ldq2 = N3 + (N3 & 1);
    const int ns = std::max(1, ksplit);   // every block writes all partial sums (tasks of K = 0 write zeros)
    for (int sp = 0; sp < ns; ++sp) {
        const int k0 = (int)((long)T.nruns * sp / ns), k1 = (int)((long)T.nruns * (sp + 1) / ns);
        SepTask t3;
        t3.boff = k0;
        t3.coff = (long)sp * NP * ldq2;
        t3.aoff = (T.contig3 ? runs[b0].i : 0) + k0;
        t3.K = k1 - k0;
        S.task3.push_back(t3);
        for (long n0 = 0; n0 < N3; n0 += tn)
            for (int m0 = 0; m0 < NP; m0 += tm) S.tile3.push_back(make_int4(sp, m0, (int)n0, 0));
    }
    T.nt2 = (int)S.tile2.size();
    T.nt3 = (int)S.tile3.size();
}

constexpr size_t I4 = sizeof(int4) / sizeof(int), IL = sizeof(long) / sizeof(int), IT = sizeof(SepTask) / sizeof(int);

// Offsets of the tables for ngr groups of gsz rows (the A row list only when
// the rows of a run are not contiguous in l2); returns the ints of all.
static size_t block_layout(SepBlock& T, long ngr, int gsz, const SmallTables& S)
{
    size_t off = 0;
    auto place = [&](size_t ints) {
        const size_t o = (off + 3) / 4 * 4;
        off = o + ints;
        return o;
    };
    T.o_groups = place((size_t)ngr * gsz);
    T.o_ginfo = place((size_t)ngr * I4);
    T.o_wofs = place((size_t)ngr * IL);
    T.o_aidx2 = place(T.contig2 ? 0 : (size_t)T.nrows);
    T.o_aidx3 = place(S.aidx3.size());
    T.o_task2 = place(S.task2.size() * IT);
    T.o_task3 = place(S.task3.size() * IT);
    T.o_tile2 = place(S.tile2.size() * I4);
    T.o_tile3 = place(S.tile3.size() * I4);
    return (off + 3) / 4 * 4;
}

// buf (ints [host0, size) of the layout): the small tables, and zeros in the
// alignment gaps (never read) from host0 on.
static void fill_small(BlockTables& BT, long ngr, int gsz, const SmallTables& S)
{
    const SepBlock& T = BT.b;
    int* const buf = BT.buf.get() - BT.host0;   // indexed by layout offset, used from host0 on
    const size_t ends[9][2] = {{T.o_groups, (size_t)ngr * gsz}, {T.o_ginfo, (size_t)ngr * I4},
                               {T.o_wofs, (size_t)ngr * IL}, {T.o_aidx2, T.contig2 ? 0 : (size_t)T.nrows},
                               {T.o_aidx3, S.aidx3.size()}, {T.o_task2, S.task2.size() * IT},
                               {T.o_task3, S.task3.size() * IT}, {T.o_tile2, S.tile2.size() * I4},
                               {T.o_tile3, S.tile3.size() * I4}};
    for (int k = 0; k < 9; ++k) {
        const size_t from = std::max(BT.host0, ends[k][0] + ends[k][1]), to = k + 1 < 9 ? ends[k + 1][0] : BT.size;
        if (to > from) memset(buf + from, 0, sizeof(int) * (to - from));
    }
    auto put = [&](size_t o, const void* src, size_t bytes) {
        if (bytes) memcpy(buf + o, src, bytes);
    };
    put(T.o_aidx3, S.aidx3.data(), sizeof(int) * S.aidx3.size());
    put(T.o_task2, S.task2.data(), sizeof(SepTask) * S.task2.size());
    put(T.o_task3, S.task3.data(), sizeof(SepTask) * S.task3.size());
    put(T.o_tile2, S.tile2.data(), sizeof(int4) * S.tile2.size());
    put(T.o_tile3, S.tile3.data(), sizeof(int4) * S.tile3.size());
}

// NP basis columns, ld = padded leading dimension of U, M and Y rows; tm x tn
// GEMM tiles; the l1 contraction's K (the block's runs) cut into up to ksplit
// tasks, task s writing its own partial sum at s NP ldq2 (ldq2: the leading
// dimension of its rows).  Host threads form the row groups over ranges of
// the grouping index (runs, or l2 values balanced by rows), concatenated in
// order: the same tables as one thread.  dev_groups: groups sharing l2 are
// left to the device (sep_groups_l2), which forms the same ones.
void build_block(const PointSet& L, const std::vector<Run>& runs, int b0, int b1, int NP, int ld, int gsz, int tm,
                 int tn, int ksplit, bool by_l2, bool dev_groups, BlockTables& BT)
{
    SepBlock& T = BT.b;
    const int kstep = std::max(1, L.kstep);
    T.row0 = runs[b0].r0;
    T.nrows = runs[b1 - 1].r1 - T.row0;
    T.nruns = b1 - b0;
    T.jspan = 0;
    // per thread: its groups (row indices, gsz per group, -1 padding), their
    // l3 start and count, and their staged-weight sizes
    struct Part {
        std::vector<int> groups;
        std::vector<int4> ginfo;
        std::vector<long> wsz;
        int jspan = 0;
    };
    const int nth = std::max(1, std::min(host_threads(), T.nruns));
    std::vector<long> rstart(T.nruns + 1);
    for (int k = 0; k < T.nruns; ++k) rstart[k] = runs[b0 + k].r0 - T.row0;
    rstart[T.nruns] = T.nrows;
    // contiguous rows: each run (l1 = i) holds l2 = i, i + 1, ... in order
    std::vector<char> rc(nth, 1);
    run_threads(nth, [&](int t) {
        for (int ru = b0 + (int)((long)T.nruns * t / nth); ru < b0 + (int)((long)T.nruns * (t + 1) / nth); ++ru)
            for (int r = runs[ru].r0; r < runs[ru].r1; ++r)
                if (L.rows[r].j != L.rows[runs[ru].r0].j + (r - runs[ru].r0)) { rc[t] = 0; break; }
    });
    T.contig2 = true;
    for (char x : rc) T.contig2 = T.contig2 && x;
    T.contig3 = true;
    for (int lt = 0; lt < T.nruns; ++lt)
        if (runs[b0 + lt].i != runs[b0].i + lt || L.rows[runs[b0 + lt].r0].j != runs[b0 + lt].i) T.contig3 = false;
    // groups by l2 when asked for, the rows of (i, j) are at r0(i) + j - i, and
    // the block spans enough l1 values (256) that the groups of one l2 and
    // parity are mostly full (a block of few l1 would give groups of 1 or 2 rows)
    by_l2 = by_l2 && kstep == 2 && T.contig2 && T.contig3 && T.nruns >= 256;
    T.swap = by_l2;
    const int ia = runs[b0].i, ib = runs[b1 - 1].i;
    // per l1 of the block: its first row and its last l2 (small tables, so the
    // l2-major walk below does not touch the row table)
    std::vector<int> r0_of, jl_of;
    int jtop = ia;
    if (by_l2) {
        r0_of.resize(ib - ia + 1);
        jl_of.resize(ib - ia + 1);
        for (int i = ia; i <= ib; ++i) {
            r0_of[i - ia] = runs[b0 + i - ia].r0;
            jl_of[i - ia] = L.rows[runs[b0 + i - ia].r1 - 1].j;
            jtop = std::max(jtop, jl_of[i - ia]);
        }
    }
    SmallTables S;
    small_tables(L, runs, b0, NP, ld, tm, tn, ksplit, T, S);
    if (by_l2 && dev_groups && gsz == 32) {
        // group counts per (l2, parity) class: rows at l2 = j of the l1 values
        // i of that parity with i <= j <= jl_of(i) (difference arrays over j)
        std::vector<int> d[2];
        for (auto& x : d) x.assign(jtop - ia + 2, 0);
        for (int i = ia; i <= ib; ++i) {
            d[i & 1][i - ia] += 1;
            d[i & 1][jl_of[i - ia] - ia + 1] -= 1;
        }
        BT.gstart.assign(2 * (jtop - ia + 1) + 1, 0);
        int c[2] = {0, 0};
        for (int j = ia; j <= jtop; ++j)
            for (int par = 0; par < 2; ++par) {
                c[par] += d[par][j - ia];
                const int cls = 2 * (j - ia) + par;
                BT.gstart[cls + 1] = BT.gstart[cls] + (c[par] + gsz - 1) / gsz;
            }
        const long ngr = BT.gstart.back();
        T.ngroups = (int)ngr;
        BT.size = block_layout(T, ngr, gsz, S);
        BT.dev = true;
        BT.host0 = T.o_aidx2;
        BT.ia = ia;
        BT.ib = ib;
        BT.jtop = jtop;
        BT.r0_of = std::move(r0_of);
        BT.jl_of = std::move(jl_of);
        BT.buf.reset(new int[BT.size - BT.host0]);
        fill_small(BT, ngr, gsz, S);
        return;   // jspan, wsize: from the device
    }
    auto last_j = [&](int i) { return jl_of[i - ia]; };
    // per thread: a range of the grouping index (runs, or l2 values balanced by rows)
    std::vector<int> cut(nth + 1);
    if (!by_l2) {
        for (int t = 0; t <= nth; ++t)
            cut[t] = (int)(std::lower_bound(rstart.begin(), rstart.end(), (long)T.nrows * t / nth) - rstart.begin());
        cut[nth] = T.nruns;
    } else {
        // rows at l2 = j: the i in [ia, min(j, ib)] whose runs reach j (a
        // difference array over j: +1 at i, -1 past last_j(i)); jstart = their
        // running sum, cut at equal shares
        std::vector<long> cnt(jtop - ia + 2, 0);
        for (int i = ia; i <= ib; ++i) {
            cnt[i - ia] += 1;
            cnt[last_j(i) - ia + 1] -= 1;
        }
        std::vector<long> jstart(jtop - ia + 2, 0);
        long run = 0, acc = 0;
        for (int j = ia; j <= jtop; ++j) {
            run += cnt[j - ia];
            jstart[j - ia] = acc;
            acc += run;
        }
        jstart[jtop - ia + 1] = acc;
        for (int t = 0; t <= nth; ++t)
            cut[t] = ia + (int)(std::lower_bound(jstart.begin(), jstart.end(), acc * t / nth) - jstart.begin());
        cut[0] = ia;
        cut[nth] = jtop + 1;
    }
    std::vector<Part> parts(nth);
    run_threads(nth, [&](int t) {
        Part& P = parts[t];
        // room for the part's rows and one partial group per class and grouping index
        const long prow = by_l2 ? T.nrows / nth + 1 : rstart[cut[t + 1]] - rstart[cut[t]];
        const long pidx = std::max(0, cut[t + 1] - cut[t]);
        P.groups.reserve(prow + (long)gsz * kstep * pidx);
        P.ginfo.reserve(prow / gsz + kstep * pidx + 1);
        P.wsz.reserve(prow / gsz + kstep * pidx + 1);
        std::vector<std::vector<int>> cls(kstep);
        auto flush = [&](std::vector<int>& c) {
            if (c.empty()) return;
            int kmin = INT_MAX, kmax = INT_MIN, jhi = INT_MIN;
            for (int g = 0; g < gsz; ++g) {
                P.groups.push_back(g < (int)c.size() ? c[g] : -1);
                if (g < (int)c.size() && L.rows[c[g]].nk > 0) {
                    const RowDesc& d = L.rows[c[g]];
                    kmin = std::min(kmin, d.k0);
                    kmax = std::max(kmax, d.k0 + kstep * (d.nk - 1));
                }
                if (g < (int)c.size()) {
                    // span of the leg that varies within the group
                    const RowDesc &d = L.rows[c[g]], &d0 = L.rows[c[0]];
                    P.jspan = std::max(P.jspan, by_l2 ? d.i - d0.i : d.j - d0.j);
                    jhi = std::max(jhi, by_l2 ? d.i : d.j);
                }
            }
            const int nkk = kmax >= kmin ? (kmax - kmin) / kstep + 1 : 0;
            // the shared leg, and the first row's and the largest other leg
            const RowDesc& d0 = L.rows[c[0]];
            const int gi = by_l2 ? d0.j : d0.i, jlo = by_l2 ? d0.i : d0.j;
            P.ginfo.push_back(make_int4(nkk ? kmin : 0, nkk, gi, jlo | (jhi << 16)));
            P.wsz.push_back((long)gsz * 16 * ((nkk + 15) / 16));
            c.clear();
        };
        if (!by_l2) {
            for (int ru = b0 + cut[t]; ru < b0 + cut[t + 1]; ++ru) {
                const Run& R = runs[ru];
                for (int r = R.r0; r < R.r1; ++r) {
                    std::vector<int>& c = cls[((L.rows[r].k0 % kstep) + kstep) % kstep];
                    c.push_back(r);
                    if ((int)c.size() == gsz) flush(c);
                }
                for (auto& c : cls) flush(c);
            }
        } else {
            // l2 = j: the rows (i, j), i of one parity at a time (one l3 grid
            // per parity), i ascending
            for (int j = cut[t]; j < cut[t + 1]; ++j) {
                for (int par = 0; par < 2; ++par) {
                    std::vector<int>& c = cls[par];
                    for (int i = ia + ((ia & 1) != par); i <= std::min(j, ib); i += 2) {
                        if (j > jl_of[i - ia]) continue;
                        c.push_back(r0_of[i - ia] + (j - i));
                        if ((int)c.size() == gsz) flush(c);
                    }
                    flush(c);
                }
            }
        }
    });
    // concatenate in order; wofs = running sum of the weight sizes
    std::vector<long> gofs(nth + 1, 0);
    for (int t = 0; t < nth; ++t) {
        gofs[t + 1] = gofs[t] + (long)parts[t].ginfo.size();
        T.jspan = std::max(T.jspan, parts[t].jspan);
    }
    const long ngr = gofs[nth];
    std::vector<long> wofs(ngr);
    long wsize = 0;
    for (int t = 0; t < nth; ++t)
        for (size_t g = 0; g < parts[t].wsz.size(); ++g) {
            wofs[gofs[t] + g] = wsize;
            wsize += parts[t].wsz[g];
        }
    T.ngroups = (int)ngr;
    T.wsize = wsize;
    BT.size = block_layout(T, ngr, gsz, S);
    BT.host0 = 0;
    BT.buf.reset(new int[BT.size]);
    int* const buf = BT.buf.get();
    fill_small(BT, ngr, gsz, S);
    // the large tables in parallel, from each thread's part
    run_threads(nth, [&](int t) {
        const Part& P = parts[t];
        if (!P.groups.empty())
            memcpy(buf + T.o_groups + (size_t)gofs[t] * gsz, P.groups.data(), sizeof(int) * P.groups.size());
        if (!P.ginfo.empty())
            memcpy(buf + T.o_ginfo + (size_t)gofs[t] * I4, P.ginfo.data(), sizeof(int4) * P.ginfo.size());
        const long g0 = gofs[t], g1 = gofs[t + 1];
        if (g1 > g0) memcpy(buf + T.o_wofs + (size_t)g0 * IL, wofs.data() + g0, sizeof(long) * (g1 - g0));
        if (!T.contig2)
            for (int ru = b0 + cut[t]; ru < b0 + cut[t + 1]; ++ru)
                for (int r = runs[ru].r0; r < runs[ru].r1; ++r) buf[T.o_aidx2 + (r - T.row0)] = L.rows[r].j;
    });
}

// blocks for ntiles tiles in groups of GS dealt to 8 XCDs
unsigned sep_blocks(long ntiles, int GS)
{
    const long groups = (ntiles + GS - 1) / GS;
    return (unsigned)(8 * GS * ((groups + 7) / 8));
}

}  // namespace

// Tables of the rows [row_lo, row_hi) of L, uploaded to c->sg_tab (kept for
// the next call on the same point set, generation and rows), and the row
// pointers in c->sg_rp.
static int separable_plan(mcmb_ctx* c, PointSet& L, int row_lo, int row_hi, int NP, int ld, int gsz, int tm,
                          int tn, int ksplit, bool& rowptr_ready)
{
    SepPlan& P = c->sep;
    const char* brenv = getenv("MCMB_SEP_BLOCK_ROWS");
    const long block_rows = brenv ? atol(brenv) : 0;
    // l-space rows are grouped by l2 (less l3-range padding per group than
    // groups sharing l1: 3.6% against 11% at lmax 2500); MCMB_SEP_GROUPS=l1
    // keeps the groups of one l1
    const char* genv = getenv("MCMB_SEP_GROUPS");
    const bool by_l2 = !L.kspace && !L.uniform && L.kstep == 2 && !(genv && !strcmp(genv, "l1"));
    const bool same_set = P.ps == &L && P.gen == L.generation;
    rowptr_ready = same_set && c->sg_rp.p;
    if (same_set && P.row_lo == row_lo && P.row_hi == row_hi && P.gsz == gsz && P.nb == NP && P.tm == tm &&
        P.ksplit == ksplit && P.block_rows == block_rows && P.by_l2 == by_l2)
        return MCMB_OK;
    setup_part(c, "before the row tables");
    P = SepPlan();
    // runs of rows with equal l1, and blocks of runs whose stage outputs fit
    // the caps (M: 16 GiB, Y: 8 GiB)
    // run starts found by host threads over ranges of rows, in order
    std::vector<Run> runs;
    {
        const int nth = std::max(1, std::min(host_threads(), (row_hi - row_lo) / 4096 + 1));
        std::vector<std::vector<int>> starts(nth);
        run_threads(nth, [&](int t) {
            const int r0 = row_lo + (int)((long)(row_hi - row_lo) * t / nth);
            const int r1 = row_lo + (int)((long)(row_hi - row_lo) * (t + 1) / nth);
            for (int r = r0; r < r1; ++r)
                if (r == row_lo || L.rows[r].i != L.rows[r - 1].i) starts[t].push_back(r);
        });
        std::vector<int> st;
        for (auto& v : starts) st.insert(st.end(), v.begin(), v.end());
        for (size_t k = 0; k < st.size(); ++k)
            runs.push_back({st[k], k + 1 < st.size() ? st[k + 1] : row_hi, L.rows[st[k]].i});
    }
    const long NPL = (long)NP * ld;
    long cap_rows = std::max<long>(1, (16l << 30) / ((long)ld * 8));
    const long cap_runs = std::max<long>(1, (8l << 30) / (NPL * 8));
    // MCMB_SEP_BLOCK_ROWS lowers the row cap (a test of the multi-block path
    // on small problems)
    if (const char* e = getenv("MCMB_SEP_BLOCK_ROWS")) cap_rows = std::max(1l, std::min(cap_rows, atol(e)));
    std::vector<std::pair<int, int>> blocks;
    for (int b0 = 0; b0 < (int)runs.size();) {
        int b1 = b0 + 1;
        long nr = runs[b0].r1 - runs[b0].r0;
        while (b1 < (int)runs.size() && b1 - b0 < cap_runs && nr + (runs[b1].r1 - runs[b1].r0) <= cap_rows) {
            nr += runs[b1].r1 - runs[b1].r0;
            ++b1;
        }
        blocks.push_back({b0, b1});
        P.max_rows = std::max(P.max_rows, nr);
        P.max_runs = std::max<long>(P.max_runs, b1 - b0);
        b0 = b1;
    }
    setup_part(c, "row runs and blocks");
    // row groups sharing l2 are formed on the device unless MCMB_SEP_TABLES=host
    // (the host walk, same tables)
    const char* tenv = getenv("MCMB_SEP_TABLES");
    const bool dev_groups = !(tenv && !strcmp(tenv, "host"));
    std::vector<BlockTables> tabs(blocks.size());
    for (size_t b = 0; b < blocks.size(); ++b) {
        build_block(L, runs, blocks[b].first, blocks[b].second, NP, ld, gsz, tm, tn, ksplit, by_l2, dev_groups, tabs[b]);
        P.max_tab = std::max(P.max_tab, tabs[b].size);
    }
    P.max_tab = (P.max_tab + 3) / 4 * 4;   // every block's tables 16-byte aligned
    setup_part(c, "row groups and tables (host)");
    MCMB_TRY(c->sg_tab.alloc(sizeof(int) * P.max_tab * std::max<size_t>(1, blocks.size())));
    setup_part(c, "table allocation");
    // device-formed blocks: inputs (per l1 first row and last l2, class
    // starts) and results (jspan, weight total) in one buffer per block
    std::vector<DevBuf> aux(blocks.size());
    bool any_dev = false;
    for (size_t b = 0; b < blocks.size(); ++b) {
        BlockTables& BT = tabs[b];
        int* const dst = c->sg_tab.as<int>() + b * P.max_tab;
        MCMB_TRY(copy_to_device(c, dst + BT.host0, BT.buf.get(), sizeof(int) * (BT.size - BT.host0)));
        if (!BT.dev) continue;
        any_dev = true;
        const size_t nl1 = BT.r0_of.size(), ncl = BT.gstart.size();
        // [jspan, pad, total (long)] then r0_of, jl_of, gstart
        std::vector<int> in(4 + 2 * nl1 + ncl, 0);
        std::copy(BT.r0_of.begin(), BT.r0_of.end(), in.begin() + 4);
        std::copy(BT.jl_of.begin(), BT.jl_of.end(), in.begin() + 4 + nl1);
        std::copy(BT.gstart.begin(), BT.gstart.end(), in.begin() + 4 + 2 * nl1);
        MCMB_TRY(aux[b].alloc(sizeof(int) * in.size()));
        MCMB_HIP(hipMemcpyAsync(aux[b].p, in.data(), sizeof(int) * in.size(), hipMemcpyHostToDevice, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));   // in is pageable and goes out of scope
        // alignment gaps of the device part: zero
        MCMB_HIP(hipMemsetAsync(dst, 0, sizeof(int) * BT.host0, c->stream));
        const SepBlock& T = BT.b;
        int* const ax = aux[b].as<int>();
        const unsigned ncls = (unsigned)(ncl - 1);
        MCMB_TRY(launch_guard("row groups", ncls, 64));
        hipLaunchKernelGGL((sep_groups_l2<32>), dim3(ncls), dim3(64), 0, c->stream, (const int*)(ax + 4),
                           (const int*)(ax + 4 + nl1), (const int*)(ax + 4 + 2 * nl1), BT.ia, BT.ib,
                           (const RowDesc*)L.d_rows.as<RowDesc>(), L.kstep, dst + T.o_groups,
                           reinterpret_cast<int4*>(dst + T.o_ginfo), reinterpret_cast<long*>(dst + T.o_wofs), ax);
        hipLaunchKernelGGL(sep_wofs_scan, dim3(1), dim3(1024), 0, c->stream, reinterpret_cast<long*>(dst + T.o_wofs),
                           (long)T.ngroups, reinterpret_cast<long*>(ax + 2));
        MCMB_HIP(hipGetLastError());
    }
    if (any_dev) {
        MCMB_HIP(hipStreamSynchronize(c->stream));
        for (size_t b = 0; b < blocks.size(); ++b) {
            if (!tabs[b].dev) continue;
            int res[4];
            MCMB_HIP(hipMemcpy(res, aux[b].p, sizeof(res), hipMemcpyDeviceToHost));
            tabs[b].b.jspan = res[0];
            long tot;
            memcpy(&tot, res + 2, sizeof(long));
            tabs[b].b.wsize = tot;
        }
    }
    for (size_t b = 0; b < blocks.size(); ++b) {
        P.max_wsize = std::max(P.max_wsize, tabs[b].b.wsize);
        P.blocks.push_back(tabs[b].b);
    }
    setup_part(c, any_dev ? "row groups (device) and table upload" : "table upload");
    P.ps = &L;
    P.gen = L.generation;
    P.row_lo = row_lo;
    P.row_hi = row_hi;
    P.gsz = gsz;
    P.nb = NP;
    P.tm = tm;
    P.ksplit = ksplit;
    P.block_rows = block_rows;
    P.by_l2 = by_l2;
    return MCMB_OK;
}

// Stage-1 geometry for NB basis columns: on FP64 MFMA over 32-row groups (1,
// 2, 4 or 8 waves of 4, 6 or 8 column tiles: NB <= 1024), else on the VALU
// over 16-row groups; MCMB_SEP_L3=valu selects the VALU kernel.  The fewest
// padding tiles, then the fewest waves (a workgroup of 4 or 8 waves loads the
// 4 SIMDs of a CU evenly; 5 waves of 6 tiles at 465 pairs ran 35% slower than
// 4 waves of 8).  ldu: the row width of the basis table the kernel reads (its
// columns, or whole 64-column blocks for the VALU kernel; padding 0).
struct L3Shape {
    int cpw = 0, nwv = 0;   // MFMA column tiles per wave, waves
    bool mfma = false;
    int gsz = SR_G, ldu = 0;
};
static L3Shape sep_l3_shape(int NB)
{
    L3Shape s;
    const char* l3env = getenv("MCMB_SEP_L3");
    const int nct = (NB + 15) / 16;   // 16-column tiles
    for (int w : {1, 2, 4, 8})
        for (int cc : {8, 6, 4})
            if (cc * w >= nct && (!s.cpw || cc * w < s.cpw * s.nwv)) s.cpw = cc, s.nwv = w;
    s.mfma = s.cpw > 0 && !(l3env && !strcmp(l3env, "valu"));
    s.gsz = s.mfma ? SR_GM : SR_G;
    s.ldu = s.mfma ? 16 * s.cpw * s.nwv : (NB + 63) / 64 * 64;
    return s;
}

// Relative column norm below which a pivoted Gram-Schmidt residual counts as
// zero.  The Legendre bases leave a gap of ~13 decades there: at lmax 2500 and
// R = 30 the 59th pivot is 1.3e-2 of the first and the 60th 4.9e-16.
constexpr double SEP_RANK_TOL = 1e-12;
// The cut is taken only across a gap: the last kept pivot must be at least
// SEP_RANK_GAP times the first dropped one (column norms).  A basis whose
// residual norms decay gradually has no numerical rank worth cutting at, and
// its pairs are contracted as they are.
constexpr double SEP_RANK_GAP = 1e6;

// Rank-revealing factorisation U = P Ct of the pair products of L into
// c->sepc, by pivoted modified Gram-Schmidt over the NP columns (one launch
// per step; the rank is read back every 8 steps).  r = 0 when it needs more
// than NP / 2 steps (too little saving to pay for the expansion) or the
// products vanish.  Kept per point set; the expansion tables are built with
// it.
// The basis a compression is made of: its cache key (the owner's address and
// generation) and the table q[l][col] (row stride ldq, nl rows).
struct BasisRef {
    const void* key;
    long gen;
    const double* q;
    int ldq;
};

static int separable_compress(mcmb_ctx* c, SepComp& S, const BasisRef& Bq, int nl, int NP, int ld,
                              const std::vector<int2>& pxy)
{
    if (S.ps == Bq.key && S.gen == Bq.gen && S.NP == NP && S.nl == nl) return MCMB_OK;
    S.ps = nullptr;
    S.r = -1;
    S.ldu = 0;
    const int nlp = nl + (nl & 1);
    const int amax = NP / 2;
    DevBuf uc, nrm, dpxy, dpiv;
    MCMB_TRY(uc.alloc(sizeof(double) * (size_t)NP * nlp));
    MCMB_TRY(dpiv.alloc(sizeof(double) * ((size_t)amax + 2)));
    MCMB_TRY(nrm.alloc(sizeof(double) * (2 * (size_t)NP + 2)));
    MCMB_TRY(dpxy.alloc(sizeof(int2) * (size_t)NP));
    MCMB_TRY(S.pc.alloc(sizeof(double) * (size_t)std::max(1, amax) * nlp));
    MCMB_TRY(S.ct.alloc(sizeof(double) * (size_t)std::max(1, amax) * ld));
    double* const dn = nrm.as<double>();
    double* const n0 = dn + 2 * NP;
    int* const rank = reinterpret_cast<int*>(n0 + 1);
    MCMB_HIP(hipMemcpyAsync(dpxy.p, pxy.data(), sizeof(int2) * (size_t)NP, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemsetAsync(S.ct.p, 0, S.ct.bytes, c->stream));
    MCMB_HIP(hipMemsetAsync(rank, 0xff, sizeof(int), c->stream));   // -1: running
    const long e = (long)NP * nlp;
    MCMB_TRY(launch_guard("pair columns", (unsigned long long)((e + 255) / 256), 256));
    hipLaunchKernelGGL(sep_ucols, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, c->stream, Bq.q, Bq.ldq, nl, nlp,
                       dpxy.as<int2>(), NP, uc.as<double>());
    const unsigned ncb = (unsigned)((NP + 3) / 4);
    hipLaunchKernelGGL(sep_mgs_norms, dim3(ncb), dim3(256), 0, c->stream, (const double*)uc.as<double>(), nlp, NP,
                       dn);
    MCMB_HIP(hipGetLastError());
    int hr = -1;
    for (int a0 = 0; a0 <= amax && hr == -1; a0 += 8) {
        for (int a = a0; a < a0 + 8 && a <= amax; ++a)
            hipLaunchKernelGGL(sep_mgs_step, dim3(ncb), dim3(256), 0, c->stream, uc.as<double>(), nlp, NP,
                               (const double*)(dn + (a & 1) * NP), dn + ((a + 1) & 1) * NP, a, amax,
                               SEP_RANK_TOL * SEP_RANK_TOL, n0, rank, dpiv.as<double>(), S.pc.as<double>(),
                               S.ct.as<double>(), ld);
        MCMB_HIP(hipGetLastError());
        MCMB_HIP(hipMemcpyAsync(&hr, rank, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
    }
    int r = hr > 0 ? hr : 0;
    if (r > 0) {
        // pivots r - 1 (last kept) and r (first dropped), squared norms
        double pv[2] = {0.0, 0.0};
        MCMB_HIP(hipMemcpyAsync(pv, dpiv.as<double>() + (r - 1), 2 * sizeof(double), hipMemcpyDeviceToHost,
                                c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        S.gap = pv[1] > 0.0 ? sqrt(pv[0] / pv[1]) : INFINITY;
        if (!(pv[0] >= SEP_RANK_GAP * SEP_RANK_GAP * pv[1])) r = 0;
    }
    if (r > 0) {
        // expansion: T1[(a,b)][z] = sum_g H[g][(a,b)] Ct[g][z]; per a,
        // Y2s[a][{y,z}] = sum_b Ct[b][y] T1[(a,b)][z] (y <= z); G[x][{y,z}] =
        // sum_a Ct[a][x] Y2s[a][{y,z}] (tile groups share a B panel, as in the
        // contractions)
        const long ldh = (long)r * r + ((r * r) & 1);
        const long NPP = (long)NP * (NP + 1) / 2;
        std::vector<SepTask> t1(1), t2(r), t3(1);
        std::vector<int4> g1, g2, g3;
        t1[0] = SepTask{0, 0, 0, r};
        for (int n0c = 0; n0c < NP; n0c += SG_TN)
            for (int m0 = 0; m0 < r * r; m0 += SG_TM) g1.push_back(make_int4(0, m0, n0c, 0));
        // Y2[a] is symmetric (H is): its upper triangle goes straight to pair
        // order, and tiles wholly below the diagonal are skipped
        const long ldp = NPP + (NPP & 1);
        for (int a = 0; a < r; ++a) {
            t2[a] = SepTask{(long)a * r, (long)a * ldp, 0, r};
            for (int n0c = 0; n0c < NP; n0c += SG_TN)
                for (int m0 = 0; m0 < NP && m0 < n0c + SG_TN; m0 += SG_TM) g2.push_back(make_int4(a, m0, n0c, 0));
        }
        // G[x][{y,z}] only where x <= y (the assembly reads sorted triples):
        // tiles whose rows all lie above the pair columns' last y are skipped
        t3[0] = SepTask{0, 0, 0, r};
        std::vector<int> ylast((NPP + SG_TN - 1) / SG_TN);
        for (int y = 0; y < NP; ++y)
            for (int z = y; z < NP; ++z) {
                const long pidx = pair_index(y, z, NP);
                ylast[pidx / SG_TN] = y;
            }
        for (long n0c = 0; n0c < NPP; n0c += SG_TN)
            for (int m0 = 0; m0 < NP && m0 <= ylast[n0c / SG_TN]; m0 += SG_TM)
                g3.push_back(make_int4(0, m0, (int)n0c, 0));
        std::vector<int> buf;
        S.o_task[0] = append(buf, t1);
        S.o_tile[0] = append(buf, g1);
        S.o_task[1] = append(buf, t2);
        S.o_tile[1] = append(buf, g2);
        S.o_task[2] = append(buf, t3);
        S.o_tile[2] = append(buf, g3);
        S.nt[0] = (int)g1.size();
        S.nt[1] = (int)g2.size();
        S.nt[2] = (int)g3.size();
        S.gs[0] = (r * r + SG_TM - 1) / SG_TM;
        S.gs[1] = S.gs[2] = (NP + SG_TM - 1) / SG_TM;
        MCMB_TRY(S.tab.alloc(sizeof(int) * buf.size()));
        MCMB_HIP(hipMemcpy(S.tab.p, buf.data(), sizeof(int) * buf.size(), hipMemcpyHostToDevice));
        MCMB_TRY(S.h.alloc(sizeof(double) * (size_t)r * ldh));
    }
    S.ps = Bq.key;
    S.gen = Bq.gen;
    S.NP = NP;
    S.nl = nl;
    S.nlp = nlp;
    S.r = r;
    return MCMB_OK;
}

// The P table [l][a] of the compressed basis with row width ldu (the stage-1
// kernel's columns), rebuilt when the width changes.
static int separable_ptab(mcmb_ctx* c, int ldu)
{
    SepComp& S = c->sepc;
    if (S.ldu == ldu) return MCMB_OK;
    // SR_UPAD zero rows after the last l (read by the fused l3 contraction)
    MCMB_TRY(S.pt.alloc(sizeof(double) * (size_t)(S.nl + SR_UPAD) * ldu));
    const long e = (long)(S.nl + SR_UPAD) * ldu;
    MCMB_TRY(launch_guard("compressed basis table", (unsigned long long)((e + 255) / 256), 256));
    hipLaunchKernelGGL(sep_ptab, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, c->stream,
                       (const double*)S.pc.as<double>(), S.nlp, S.nl, S.r, ldu, S.pt.as<double>());
    MCMB_HIP(hipGetLastError());
    S.ldu = ldu;
    return MCMB_OK;
}

// ---- Gauss-Legendre Gram over the compressed pair basis ---------------------
//
// The GL-factorised Gram (calculate_orthol, orthogonalisation_late.c:239-336)
// is Gamma_mn = sum_s w_s / (48 pi) perm[M_{m_a n_b}(x_s)] with M_p(x) =
// sum_l (2l+1) v_l P_l(x) U[l][p] over the pair products U.  With U = P Ct
// (the compressed basis of the separable Gram), M_p(x_s) = sum_a Ct[a][p]
// Phi_a(s), Phi_a(s) = sum_l P[l][a] (2l+1) v_l P_l(x_s), so the node sum
// becomes the core H[g][a][b] = (3 / 4 pi) sum_s w_s Phi_g Phi_a Phi_b (r^3
// entries instead of a permanent per mode pair and node), expanded to pairs
// and assembled exactly as the separable Gram's core (whose assembly takes
// 1/36 of the six matchings: 36 / (48 pi) = 3 / (4 pi)).

// The compressed basis with the GL weight folded in, Pw[a][l] = P[l][a]
// (2l + 1) v_l (0 below lmin).
__global__ void sep_gl_tabs(const double* __restrict__ Pc, int nlp, int nl, int r, int lmin,
                            const double* __restrict__ v, double* __restrict__ Pw)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long)r * nlp) {
        const int l = (int)(e % nlp);
        Pw[e] = (l < nl && l >= lmin) ? Pc[e] * ((2.0 * l + 1.0) * v[l]) : 0.0;
    }
}

// Phi[a][s] (row stride npt) for a < r: thread (node s, group of GL_AG
// columns) runs the Legendre recurrence over l and accumulates its columns
// (the column group is uniform over the wave, so its table reads are scalar).
// The recurrence keeps its division by l: with rounded coefficients (2l -
// 1)/l, (l - 1)/l it ran 10% faster but moved the lmax-500 Gram by 3e-12.
constexpr int GL_AG = 4;
__global__ __launch_bounds__(256) void sep_gl_phi(const double* __restrict__ Pw, int nlp, int nl, int r,
                                                  const double* __restrict__ x, int npt, double* __restrict__ Phi)
{
    const int s = blockIdx.x * 64 + (threadIdx.x & 63);
    const int a0 = __builtin_amdgcn_readfirstlane(GL_AG * (blockIdx.y * 4 + (threadIdx.x >> 6)));
    if (a0 >= r) return;   // the whole wave
    const double xs = s < npt ? x[s] : 0.0;
    double acc[GL_AG];
    const double* pw[GL_AG];
#pragma unroll
    for (int k = 0; k < GL_AG; ++k) {
        acc[k] = 0.0;
        pw[k] = Pw + (long)min(a0 + k, r - 1) * nlp;   // columns past r repeat the last (not stored)
    }
    double p0 = 1.0, p1 = xs;   // P_{l-2}, P_{l-1}
#pragma unroll
    for (int k = 0; k < GL_AG; ++k) acc[k] = fma(pw[k][0], 1.0, acc[k]);
    if (nl > 1)
#pragma unroll
        for (int k = 0; k < GL_AG; ++k) acc[k] = fma(pw[k][1], xs, acc[k]);
    for (int l = 2; l < nl; ++l) {
        const double pl = ((2.0 * l - 1.0) * xs * p1 - (l - 1.0) * p0) / l;
        p0 = p1;
        p1 = pl;
#pragma unroll
        for (int k = 0; k < GL_AG; ++k) acc[k] = fma(pw[k][l], pl, acc[k]);
    }
    if (s < npt)
#pragma unroll
        for (int k = 0; k < GL_AG; ++k)
            if (a0 + k < r) Phi[(long)(a0 + k) * npt + s] = acc[k];
}

// H[g][a r + b] = H[g][b r + a] = scale sum_s w_s Phi_g(s) Phi_a(s) Phi_b(s),
// one workgroup per a <= b: t = w Phi_a Phi_b in LDS chunks of GL_SC nodes,
// wave q sums the columns g = q, q + 4, ... over its lanes (fixed order).
// MIX (two fields): Phi_g from PhiG (rg columns) and the entries go to
// H[b][g r + a] = H[a][g r + b] (the paired field's index first).
constexpr int GL_SC = 1024;
constexpr int GL_GW = 16;   // columns g per wave (r, rg <= 64)
template <bool MIX>
__global__ __launch_bounds__(256) void sep_gl_core(const double* __restrict__ Phi, const double* __restrict__ PhiG,
                                                   int rg, int npt, const double* __restrict__ w, int r,
                                                   double scale, double* __restrict__ H, long ldh)
{
    __shared__ double ts[GL_SC];
    int a = 0, rem = blockIdx.x;   // pair (a, b), a <= b, row-major upper triangle
    while (rem >= r - a) {
        rem -= r - a;
        ++a;
    }
    const int b = a + rem;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    double acc[GL_GW];
#pragma unroll
    for (int k = 0; k < GL_GW; ++k) acc[k] = 0.0;
    for (int s0 = 0; s0 < npt; s0 += GL_SC) {
        __syncthreads();
        for (int e = threadIdx.x; e < GL_SC; e += 256) {
            const int s = s0 + e;
            ts[e] = s < npt ? w[s] * Phi[(long)a * npt + s] * Phi[(long)b * npt + s] : 0.0;
        }
        __syncthreads();
        const int ns = min(GL_SC, npt - s0);
#pragma unroll
        for (int k = 0; k < GL_GW; ++k) {
            const int g = q + 4 * k;
            if (g < rg)
                for (int e = lane; e < ns; e += 64) acc[k] = fma(PhiG[(long)g * npt + s0 + e], ts[e], acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < GL_GW; ++k) {
        const int g = q + 4 * k;
        double t = acc[k];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) t += __shfl_xor(t, o);
        if (g < rg && lane == 0) {
            if (MIX) {
                H[(long)b * ldh + (long)g * r + a] = scale * t;
                H[(long)a * ldh + (long)g * r + b] = scale * t;
            } else {
                H[(long)g * ldh + (long)a * r + b] = scale * t;
                H[(long)g * ldh + (long)b * r + a] = scale * t;
            }
        }
    }
}

// One field of a compressed GL Gram whose three legs are that field: its
// basis (the first R columns of q[l][col], nl rows from l = 0), per-leg
// weight v, lmin, the mode table (legs packed 8 bits each) and the
// compression cache kept for that basis.
struct GlField3 {
    BasisRef B;
    int nl, R, lmin;
    const double* d_v;
    int n;
    const int* d_modes;
    SepComp* S;
};

static int gl_compressed_run(mcmb_ctx* c, const GlField3& F, int npt, const double* x, const double* w,
                             double* d_gamma, bool* used);

int gram_gl_compressed(mcmb_ctx* c, int npt, const double* x, const double* w, double* d_gamma, bool* used)
{
    *used = false;
    PointSet& L = c->ls;
    if (!gram_separable_supported(L, MCMB_FP64)) return MCMB_OK;
    const GlField3 F{BasisRef{L.basis_key(), L.basis_generation(), L.d_q.as<double>(), L.ldq},
                     (int)(L.q.size() / (size_t)L.ldq), L.R, L.lmin, L.d_v.as<double>(), L.n,
                     L.d_modes.as<int>(), &c->sepc};
    return gl_compressed_run(c, F, npt, x, w, d_gamma, used);
}

int gram_gl_eee_compressed(mcmb_ctx* c, int R, int npt, const double* x, const double* w, double* d_gamma,
                           bool* used)
{
    *used = false;
    PolSet& E = c->es;
    if (R < 1 || R > 255 || R * (R + 1) / 2 > SR_NPMAX || E.lmax >= 32768) return MCMB_OK;
    const GlField3 F{BasisRef{&c->es, E.generation, E.d_q.as<double>(), E.ldq}, E.lmax + 1, R, E.lmin,
                     E.d_v.as<double>(), E.n, E.d_modes[2].as<int>(), &c->sepc_e};
    return gl_compressed_run(c, F, npt, x, w, d_gamma, used);
}

static int gl_compressed_run(mcmb_ctx* c, const GlField3& F, int npt, const double* x, const double* w,
                             double* d_gamma, bool* used)
{
    *used = false;
    const char* genv = getenv("MCMB_GL_PATH");
    if (genv && !strcmp(genv, "direct")) return MCMB_OK;
    const int n = F.n, R = F.R, NP = R * (R + 1) / 2;
    const int ld = NP + (NP & 1);
    const int nl = F.nl;
    std::vector<int2> pxy(NP);
    for (int a = 0; a < R; ++a)
        for (int b = a; b < R; ++b) pxy[pair_index(a, b, R)] = make_int2(a, b);
    MCMB_TRY(separable_compress(c, *F.S, F.B, nl, NP, ld, pxy));
    const SepComp& S = *F.S;
    const int r = S.r;
    if (r <= 0 || r > 4 * GL_GW) return MCMB_OK;   // full rank: the direct node sum
    const long ldh = (long)r * r + ((r * r) & 1);
    const long NPP = (long)NP * (NP + 1) / 2, ldp = NPP + (NPP & 1);
    MCMB_TRY(c->gl_xw.alloc(sizeof(double) * (2 * (size_t)npt + (size_t)r * npt + (size_t)r * S.nlp)));
    double* const dx = c->gl_xw.as<double>();
    double* const dw = dx + npt;
    double* const dphi = dw + npt;
    double* const dpw = dphi + (size_t)r * npt;
    MCMB_TRY(c->sg_M.alloc(sizeof(double) * (size_t)(ldh * ld)));
    MCMB_TRY(c->sg_Ys.alloc(sizeof(double) * (size_t)((long)r * ldp)));
    MCMB_TRY(c->sg_G.alloc(sizeof(double) * (size_t)((long)NP * ldp)));
    MCMB_HIP(hipMemcpyAsync(dx, x, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(dw, w, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    const long ntab = (long)r * S.nlp;
    hipLaunchKernelGGL(sep_gl_tabs, dim3((unsigned)((ntab + 255) / 256)), dim3(256), 0, c->stream,
                       (const double*)S.pc.as<double>(), S.nlp, nl, r, F.lmin, F.d_v, dpw);
    hipLaunchKernelGGL(sep_gl_phi, dim3((unsigned)((npt + 63) / 64), (unsigned)((r + 4 * GL_AG - 1) / (4 * GL_AG))),
                       dim3(256), 0, c->stream, (const double*)dpw, S.nlp, nl, r, (const double*)dx, npt, dphi);
    MCMB_HIP(hipMemsetAsync(S.h.p, 0, sizeof(double) * (size_t)r * ldh, c->stream));
    hipLaunchKernelGGL(sep_gl_core<false>, dim3((unsigned)(r * (r + 1) / 2)), dim3(256), 0, c->stream,
                       (const double*)dphi, (const double*)dphi, r, npt, (const double*)dw, r, 3.0 / (4.0 * M_PI),
                       S.h.as<double>(), ldh);
    MCMB_HIP(hipGetLastError());
    // the expansion and the assembly of the separable Gram
    const int* tb = S.tab.as<int>();
    double* const T1 = c->sg_M.as<double>();
    double* const Y2s = c->sg_Ys.as<double>();
    double* const G = c->sg_G.as<double>();
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(S.nt[0], S.gs[0]), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN>), dim3(sep_blocks(S.nt[0], S.gs[0])), dim3(256), 0, c->stream,
                       (const double*)S.h.as<double>(), (int)ldh, tb, (const double*)S.ct.as<double>(), (long)ld, T1,
                       (long)ld, r * r, NP, reinterpret_cast<const SepTask*>(tb + S.o_task[0]),
                       reinterpret_cast<const int4*>(tb + S.o_tile[0]), S.nt[0], S.gs[0], 0);
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(S.nt[1], S.gs[1]), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN, true>), dim3(sep_blocks(S.nt[1], S.gs[1])), dim3(256), 0,
                       c->stream, (const double*)S.ct.as<double>(), ld, tb, (const double*)T1, (long)ld, Y2s, ldp, NP,
                       NP, reinterpret_cast<const SepTask*>(tb + S.o_task[1]),
                       reinterpret_cast<const int4*>(tb + S.o_tile[1]), S.nt[1], S.gs[1], 0);
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(S.nt[2], S.gs[2]), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN>), dim3(sep_blocks(S.nt[2], S.gs[2])), dim3(256), 0, c->stream,
                       (const double*)S.ct.as<double>(), ld, tb, (const double*)Y2s, (long)ldp, G, (long)ldp, NP,
                       (int)NPP, reinterpret_cast<const SepTask*>(tb + S.o_task[2]),
                       reinterpret_cast<const int4*>(tb + S.o_tile[2]), S.nt[2], S.gs[2], 0);
    constexpr int SB = 6;   // blocks of SB x SB assembly tiles per XCD
    const int nt = (n + 31) / 32, nbk = (nt + SB - 1) / SB;
    const unsigned long long ntile = ((unsigned long long)nbk * (nbk + 1) / 2 + 7) / 8 * 8 * SB * SB;
    MCMB_TRY(launch_guard("GL assembly", ntile, 256));
    hipLaunchKernelGGL(sep_assemble_sym<SB>, dim3((unsigned)ntile), dim3(256), 0, c->stream, (const double*)G, ldp, R,
                       F.d_modes, n, nt, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));   // x and w may live on the caller's stack
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 8;
    c->nphases = 0;
    c->last_rank = r;
    *used = true;
    return MCMB_OK;
}

// The mixed polarised components TTE (pol 1) and TEE (pol 2) over the two
// compressed bases.  With Phi^T, Phi^E the compressed bases evaluated at the
// nodes as above, the field P on two legs (T for TTE, E for TEE) and S on
// the third:
//   H[b][g rP + a] = 1/(16 pi) sum_s w_s PhiP_a PhiP_b PhiS_g   (a, b < rP; g < rS)
//   T1[(g, a)][q]  = sum_b H[b][g rP + a] CtP[b][q]
//   Y[g][{p, q}]   = sum_a CtP[a][p] T1[(g, a)][q]     (symmetric: pair order)
//   K[u][{p, q}]   = sum_g CtS[g][u] Y[g][{p, q}]
// so K[u][{p, q}] = 1/(16 pi) sum_s w_s MP_p MP_q MS_u, assembled by
// sep_assemble_mix.  RT, RE: the basis columns the component's modes use.
int gram_gl_mixed_compressed(mcmb_ctx* c, int pol, int RT, int RE, int npt, const double* x, const double* w,
                             double* d_gamma, bool* used)
{
    *used = false;
    const char* genv = getenv("MCMB_GL_PATH");
    if (genv && !strcmp(genv, "direct")) return MCMB_OK;
    PointSet& L = c->ls;
    PolSet& E = c->es;
    auto fits = [](int R) { return R >= 1 && R <= 255 && R * (R + 1) / 2 <= SR_NPMAX; };
    if (!fits(RT) || !fits(RE) || L.lmax >= 32768 || E.lmax >= 32768) return MCMB_OK;
    const int NPT = RT * (RT + 1) / 2, NPE = RE * (RE + 1) / 2;
    const int ldT = NPT + (NPT & 1), ldE = NPE + (NPE & 1);
    const int nlT = (int)(L.q.size() / (size_t)L.ldq), nlE = E.lmax + 1;
    auto pairs = [](int R) {
        std::vector<int2> v(R * (R + 1) / 2);
        for (int a = 0; a < R; ++a)
            for (int b = a; b < R; ++b) v[pair_index(a, b, R)] = make_int2(a, b);
        return v;
    };
    MCMB_TRY(separable_compress(c, c->sepc_pt,
                                BasisRef{L.basis_key(), L.basis_generation(), L.d_q.as<double>(), L.ldq}, nlT, NPT,
                                ldT, pairs(RT)));
    MCMB_TRY(separable_compress(c, c->sepc_e, BasisRef{&c->es, E.generation, E.d_q.as<double>(), E.ldq}, nlE, NPE,
                                ldE, pairs(RE)));
    const SepComp& ST = c->sepc_pt;
    const SepComp& SE = c->sepc_e;
    const int rT = ST.r, rE = SE.r;
    if (rT <= 0 || rT > 4 * GL_GW || rE <= 0 || rE > 4 * GL_GW) return MCMB_OK;   // the direct node sum
    // P: the field on two legs, S: the other
    const bool tte = pol == 1;
    const SepComp& SP = tte ? ST : SE;
    const SepComp& SS = tte ? SE : ST;
    const int rP = SP.r, rS = SS.r, NPP_ = tte ? NPT : NPE, NPS = tte ? NPE : NPT;
    const int ldP = tte ? ldT : ldE, ldS = tte ? ldE : ldT;
    const long NQ = (long)NPP_ * (NPP_ + 1) / 2, ldq = NQ + (NQ & 1);   // pair pairs of P
    const long ldh = (long)rS * rP + (((long)rS * rP) & 1);
    // device work: x, w, Phi^T, Phi^E, the two weighted tables, H
    const size_t nd = 2 * (size_t)npt + (size_t)(rT + rE) * npt + (size_t)rT * ST.nlp + (size_t)rE * SE.nlp +
                      (size_t)rP * ldh;
    MCMB_TRY(c->gl_xw.alloc(sizeof(double) * nd));
    double* const dx = c->gl_xw.as<double>();
    double* const dw = dx + npt;
    double* const phT = dw + npt;
    double* const phE = phT + (size_t)rT * npt;
    double* const pwT = phE + (size_t)rE * npt;
    double* const pwE = pwT + (size_t)rT * ST.nlp;
    double* const H = pwE + (size_t)rE * SE.nlp;
    MCMB_TRY(c->sg_M.alloc(sizeof(double) * (size_t)((long)rS * rP * ldP)));
    MCMB_TRY(c->sg_Ys.alloc(sizeof(double) * (size_t)((long)rS * ldq)));
    MCMB_TRY(c->sg_G.alloc(sizeof(double) * (size_t)((long)NPS * ldq)));
    // the three GEMMs' tasks and tiles (groups: the m tiles of one B panel)
    std::vector<SepTask> t1(1), t2(rS), t3(1);
    std::vector<int4> g1, g2, g3;
    t1[0] = SepTask{0, 0, 0, rP};
    for (int n0 = 0; n0 < NPP_; n0 += SG_TN)
        for (int m0 = 0; m0 < rS * rP; m0 += SG_TM) g1.push_back(make_int4(0, m0, n0, 0));
    for (int g = 0; g < rS; ++g) {
        t2[g] = SepTask{(long)g * rP, (long)g * ldq, 0, rP};
        for (int n0 = 0; n0 < NPP_; n0 += SG_TN)
            for (int m0 = 0; m0 < NPP_ && m0 < n0 + SG_TN; m0 += SG_TM) g2.push_back(make_int4(g, m0, n0, 0));
    }
    t3[0] = SepTask{0, 0, 0, rS};
    for (long n0 = 0; n0 < NQ; n0 += SG_TN)
        for (int m0 = 0; m0 < NPS; m0 += SG_TM) g3.push_back(make_int4(0, m0, (int)n0, 0));
    std::vector<int> buf;
    const size_t ot1 = append(buf, t1), og1 = append(buf, g1), ot2 = append(buf, t2), og2 = append(buf, g2),
                 ot3 = append(buf, t3), og3 = append(buf, g3);
    MCMB_TRY(c->gl_tab.alloc(sizeof(int) * buf.size()));
    const int* tb = c->gl_tab.as<int>();
    MCMB_HIP(hipMemcpyAsync(c->gl_tab.p, buf.data(), sizeof(int) * buf.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(dx, x, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(dw, w, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    // the compressed bases at the nodes
    struct Fd {
        const SepComp* S;
        int nl, lmin;
        const double* v;
        double *pw, *ph;
    };
    const Fd fd[2] = {{&ST, nlT, L.lmin, L.d_v.as<double>(), pwT, phT}, {&SE, nlE, E.lmin, E.d_v.as<double>(), pwE, phE}};
    for (const Fd& f : fd) {
        const int r = f.S->r;
        const long ntab = (long)r * f.S->nlp;
        hipLaunchKernelGGL(sep_gl_tabs, dim3((unsigned)((ntab + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)f.S->pc.as<double>(), f.S->nlp, f.nl, r, f.lmin, f.v, f.pw);
        hipLaunchKernelGGL(sep_gl_phi, dim3((unsigned)((npt + 63) / 64), (unsigned)((r + 4 * GL_AG - 1) / (4 * GL_AG))),
                           dim3(256), 0, c->stream, (const double*)f.pw, f.S->nlp, f.nl, r, (const double*)dx, npt,
                           f.ph);
    }
    const double* const phP = tte ? phT : phE;
    const double* const phS = tte ? phE : phT;
    MCMB_HIP(hipMemsetAsync(H, 0, sizeof(double) * (size_t)rP * ldh, c->stream));
    hipLaunchKernelGGL(sep_gl_core<true>, dim3((unsigned)(rP * (rP + 1) / 2)), dim3(256), 0, c->stream, phP, phS, rS,
                       npt, (const double*)dw, rP, 1.0 / (16.0 * M_PI), H, ldh);
    MCMB_HIP(hipGetLastError());
    double* const T1 = c->sg_M.as<double>();
    double* const Y = c->sg_Ys.as<double>();
    double* const K = c->sg_G.as<double>();
    const int gs1 = (rS * rP + SG_TM - 1) / SG_TM, gs2 = (NPP_ + SG_TM - 1) / SG_TM, gs3 = (NPS + SG_TM - 1) / SG_TM;
    const int nt1 = (int)g1.size(), nt2 = (int)g2.size(), nt3 = (int)g3.size();
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(nt1, gs1), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN>), dim3(sep_blocks(nt1, gs1)), dim3(256), 0, c->stream,
                       (const double*)H, (int)ldh, tb, (const double*)SP.ct.as<double>(), (long)ldP, T1, (long)ldP,
                       rS * rP, NPP_, reinterpret_cast<const SepTask*>(tb + ot1),
                       reinterpret_cast<const int4*>(tb + og1), nt1, gs1, 0);
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(nt2, gs2), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN, true>), dim3(sep_blocks(nt2, gs2)), dim3(256), 0, c->stream,
                       (const double*)SP.ct.as<double>(), ldP, tb, (const double*)T1, (long)ldP, Y, ldq, NPP_, NPP_,
                       reinterpret_cast<const SepTask*>(tb + ot2), reinterpret_cast<const int4*>(tb + og2), nt2, gs2,
                       0);
    MCMB_TRY(launch_guard("GL expansion", sep_blocks(nt3, gs3), 256));
    hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN>), dim3(sep_blocks(nt3, gs3)), dim3(256), 0, c->stream,
                       (const double*)SS.ct.as<double>(), ldS, tb, (const double*)Y, (long)ldq, K, (long)ldq, NPS,
                       (int)NQ, reinterpret_cast<const SepTask*>(tb + ot3), reinterpret_cast<const int4*>(tb + og3),
                       nt3, gs3, 0);
    const int n = E.n;
    constexpr int SB = 6;
    const int nt = (n + 31) / 32, nbk = (nt + SB - 1) / SB;
    const unsigned long long ntile = ((unsigned long long)nbk * (nbk + 1) / 2 + 7) / 8 * 8 * SB * SB;
    MCMB_TRY(launch_guard("GL assembly", ntile, 256));
    if (tte)
        hipLaunchKernelGGL((sep_assemble_mix<1, SB>), dim3((unsigned)ntile), dim3(256), 0, c->stream,
                           (const double*)K, ldq, RT, RE, E.d_modes[0].as<int>(), n, nt, d_gamma);
    else
        hipLaunchKernelGGL((sep_assemble_mix<2, SB>), dim3((unsigned)ntile), dim3(256), 0, c->stream,
                           (const double*)K, ldq, RT, RE, E.d_modes[1].as<int>(), n, nt, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));   // x, w and the tables may live on the host stack
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 11;
    c->nphases = 0;
    c->last_rank = rP;   // the basis on two legs
    *used = true;
    return MCMB_OK;
}

int gram_separable_launch(mcmb_ctx* c, PointSet& L, int precision, double* d_gamma, int row_lo, int row_hi)
{
    if (!gram_separable_supported(L, precision))
        return fail(MCMB_ERR_ARG, "separable Gram: precision %d / basis width %d not supported", precision, L.R);
    const int n = L.n, R = L.R, NP = R * (R + 1) / 2;
    const int ld = NP + (NP & 1);          // even leading dimension of pair rows: 16-byte aligned
    const int nl = (int)(L.q.size() / (size_t)L.ldq);
    c->nphases = 0;
    // host wall time of each setup part and of the whole call
    setup_begin(c);
    auto setup_done = [&](const char* name) { setup_part(c, name); };
    if (row_hi <= row_lo) {
        MCMB_HIP(hipMemsetAsync(d_gamma, 0, sizeof(double) * (size_t)n * n, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        c->last_ms = 0.0;
        c->last_launches = 0;
        return MCMB_OK;
    }
    std::vector<int2> pxy(NP);
    for (int x = 0; x < R; ++x)
        for (int y = x; y < R; ++y) pxy[pair_index(x, y, R)] = make_int2(x, y);

    // phases: [basis compression], pair products, [weights], per block the
    // three contractions (times summed over blocks), [expansion], assembly
    int np = 0;   // phases started; phase k runs from event k to event k + 1
    const char* penv = getenv("MCMB_PHASES");
    const bool phases = !(penv && !strcmp(penv, "0"));
    auto phase = [&](const char* name) -> int {
        if (!phases) return MCMB_OK;
        c->phase_names[np] = name;
        MCMB_HIP(hipEventRecord(c->evp[np], c->stream));
        ++np;
        return MCMB_OK;
    };
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    int launches = 0;
    // The contractions run over NB basis columns: the pairs, or the r columns
    // of the compressed basis when the pair products have rank r <= NP / 2
    // (MCMB_SEP_COMPRESS=0: always the pairs).
    const char* cenv = getenv("MCMB_SEP_COMPRESS");
    SepComp& S = c->sepc;
    bool comp = false;
    if (!(cenv && !strcmp(cenv, "0"))) {
        if (!(S.ps == L.basis_key() && S.gen == L.basis_generation() && S.NP == NP && S.nl == nl)) {
            MCMB_TRY(phase("basis compression"));
            launches += 2 + NP / 2;
        }
        MCMB_TRY(separable_compress(c, c->sepc,
                                    BasisRef{L.basis_key(), L.basis_generation(), L.d_q.as<double>(), L.ldq}, nl, NP,
                                    ld, pxy));
        comp = S.r > 0;
        setup_done("basis factorisation");
    }
    const int NB = comp ? S.r : NP;
    const int ldB = NB + (NB & 1);         // even leading dimension of M, Y rows
    const L3Shape sh = sep_l3_shape(NB);
    const int cpw = sh.cpw, nwv = sh.nwv, gsz = sh.gsz, ldu = sh.ldu;
    const bool l3mfma = sh.mfma;
    const long NPL = (long)NB * ldB;       // one Y slab

    // narrow bases: 64 x 64 GEMM tiles; over a compressed basis the l1
    // contraction's K is split (its partial sums are added by the expansion)
    const bool small = NB <= SG_SMALL;
    const int tm = small ? 64 : SG_TM, tn = small ? 64 : SG_TN;
    const int ksplit = comp ? SG_KSPLIT : 1;
    bool rowptr_ready = false;
    MCMB_TRY(separable_plan(c, L, row_lo, row_hi, NB, ldB, gsz, tm, tn, ksplit, rowptr_ready));
    const SepPlan& P = c->sep;

    const long NPP = (long)NP * (NP + 1) / 2;   // unordered pairs {e, f}
    const long ldp = NPP + (NPP & 1);          // even leading dimension of Ys and G2 rows
    const long NGS = NPP * (NP + 2) / 3;       // x <= y <= z triples: NP (NP + 1) (NP + 2) / 6
    const long NBP = (long)NB * (NB + 1) / 2;  // unordered basis pairs
    const long ldq2 = NBP + (NBP & 1);
    const long ldh = (long)NB * NB + ((NB * NB) & 1);
    MCMB_TRY(c->sg_M.alloc(sizeof(double) * (size_t)std::max(P.max_rows * ldB, comp ? ldh * ld : 0l)));
    MCMB_TRY(c->sg_Y.alloc(sizeof(double) * (size_t)(P.max_runs * NPL)));
    MCMB_TRY(c->sg_Ys.alloc(sizeof(double) * (size_t)std::max(P.max_runs * ldq2, comp ? (long)NB * ldp : 0l)));
    MCMB_TRY(c->sg_G.alloc(sizeof(double) * (size_t)std::max(NP * ldp, ksplit * NB * ldq2)));
    if (!comp) MCMB_TRY(c->sg_Gs.alloc(sizeof(double) * (size_t)NGS));
    if (!comp && c->sg_off_np != NP) {
        // off[{x,y}] = first Gs element of the triples (x, y, z >= y)
        std::vector<long> off(NPP);
        long o = 0;
        for (int x = 0; x < NP; ++x)
            for (int y = x; y < NP; ++y) {
                off[pair_index(x, y, NP)] = o;
                o += NP - y;
            }
        MCMB_TRY(c->sg_off.alloc(sizeof(long) * (size_t)NPP));
        MCMB_HIP(hipMemcpy(c->sg_off.p, off.data(), sizeof(long) * (size_t)NPP, hipMemcpyHostToDevice));
        c->sg_off_np = NP;
    }

    setup_done("buffers");
    const double* dU;   // basis table [l][ldu] the contractions read
    if (comp) {
        MCMB_TRY(separable_ptab(c, ldu));
        dU = S.pt.as<double>();
    } else {
        MCMB_TRY(c->sg_U.alloc(sizeof(double) * (size_t)(nl + SR_UPAD) * ldu + sizeof(int2) * NP));
        double* const dUw = c->sg_U.as<double>();
        int2* d_pxy = reinterpret_cast<int2*>(dUw + (size_t)(nl + SR_UPAD) * ldu);
        MCMB_HIP(hipMemcpy(d_pxy, pxy.data(), sizeof(int2) * NP, hipMemcpyHostToDevice));
        MCMB_TRY(phase("pair products"));
        const long e = (long)(nl + SR_UPAD) * ldu;
        MCMB_TRY(launch_guard("pair products", (unsigned long long)((e + 255) / 256), 256));
        hipLaunchKernelGGL(sep_utab, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, c->stream, L.d_q.as<double>(),
                           L.ldq, nl, d_pxy, NP, ldu, dUw);
        MCMB_HIP(hipGetLastError());
        ++launches;
        dU = dUw;
    }

    const bool one_block = P.blocks.size() == 1;
    // l-space weights are formed where they are staged; other point sets stage
    // their walk-order weights
    const int wkind = (L.kspace || L.uniform) ? 2 : (L.h2mode == MCMB_H2_ASYMPTOTIC ? 1 : 0);
    // the l3 stage reads (2 l3 + 1) v_l3 U(l3) for the exact l-space weight,
    // which leaves that factor out (sep_weight); U itself otherwise.  Kept for
    // the same table, setup and width.
    const double* dU1 = dU;
    if (wkind == 0) {
        if (!(c->sg_u3_src == (const void*)dU && c->sg_u3_gen == L.generation && c->sg_u3_ldu == ldu &&
              c->sg_u3_nl == nl)) {
            MCMB_TRY(c->sg_U3.alloc(sizeof(double) * (size_t)(nl + SR_UPAD) * ldu));
            const long e = (long)(nl + SR_UPAD) * ldu;
            MCMB_TRY(launch_guard("l3 basis table", (unsigned long long)((e + 255) / 256), 256));
            hipLaunchKernelGGL(sep_u3tab, dim3((unsigned)((e + 255) / 256)), dim3(256), 0, c->stream, dU, ldu, nl,
                               (const double*)L.d_vp.as<double>(), c->sg_U3.as<double>());
            MCMB_HIP(hipGetLastError());
            ++launches;
            c->sg_u3_src = dU;
            c->sg_u3_gen = L.generation;
            c->sg_u3_ldu = ldu;
            c->sg_u3_nl = nl;
        }
        dU1 = c->sg_U3.as<double>();
    }
    // the MFMA l3 contraction forms l-space weights itself (1, 2 or 4 waves:
    // up to 512 columns); MCMB_SEP_WEIGHTS=staged stages them instead
    const char* wenv = getenv("MCMB_SEP_WEIGHTS");
    // (one wave reads its tables through 64-entry windows: the rows of a group
    // must span at most 96 in j, which the walk's 32-row groups of one parity
    // class do)
    int jspan = 0;
    for (const SepBlock& T : P.blocks) jspan = std::max(jspan, T.jspan);
    const bool wfused = l3mfma && SR_GM == 32 && nwv <= 4 && (nwv == 1 || cpw != 4) && wkind != 2 && L.kstep == 2 &&
                        (nwv > 1 || jspan <= 96) && !(wenv && !strcmp(wenv, "staged"));
    if (!wfused) MCMB_TRY(c->sg_wd.alloc(sizeof(double) * (size_t)std::max<long>(1, P.max_wsize)));
    // staged weights of one block's row groups
    auto wstage = [&](size_t b) -> int {
        const SepBlock& T = P.blocks[b];
        const int* tab = c->sg_tab.as<int>() + b * P.max_tab;
        MCMB_TRY(launch_guard("weight staging", (unsigned long long)T.ngroups, 256));
#define SEP_WSTAGE(KK, GG)                                                                                      \
    hipLaunchKernelGGL((sep_wstage<KK, GG>), dim3((unsigned)T.ngroups), dim3(256), 0, c->stream,                \
                       L.d_rows.as<RowDesc>(), (const long*)c->sg_rp.as<long>(), (const double*)c->sg_ws.as<double>(), \
                       tab + T.o_groups, reinterpret_cast<const int4*>(tab + T.o_ginfo),                        \
                       reinterpret_cast<const long*>(tab + T.o_wofs), L.kstep,                                   \
                       (const double*)L.d_ftab.as<double>(), (const double*)L.d_v.as<double>(),                  \
                       c->sg_wd.as<double>())
        if (gsz == SR_GM) {
            if (wkind == 0) SEP_WSTAGE(0, SR_GM);
            else if (wkind == 1) SEP_WSTAGE(1, SR_GM);
            else SEP_WSTAGE(2, SR_GM);
        } else {
            if (wkind == 0) SEP_WSTAGE(0, SR_G);
            else if (wkind == 1) SEP_WSTAGE(1, SR_G);
            else SEP_WSTAGE(2, SR_G);
        }
#undef SEP_WSTAGE
        MCMB_HIP(hipGetLastError());
        ++launches;
        return MCMB_OK;
    };
    if (!wfused) {
        MCMB_TRY(phase("weights"));
        if (wkind == 2) {
            MCMB_TRY(pointset_weights(c, L, c->sg_rp, c->sg_ws, 0, rowptr_ready));
            ++launches;
        }
        if (one_block) MCMB_TRY(wstage(0));
    }
    if (!one_block) MCMB_TRY(phase("l3, l2, l1 contractions (all blocks)"));
    const int GS = (NB + tm - 1) / tm;   // m tiles per B panel, both contractions
    for (size_t b = 0; b < P.blocks.size(); ++b) {
        const SepBlock& T = P.blocks[b];
        const int* tab = c->sg_tab.as<int>() + b * P.max_tab;
        const int4* ginfo = reinterpret_cast<const int4*>(tab + T.o_ginfo);
        const long* wofs = reinterpret_cast<const long*>(tab + T.o_wofs);
        if (!one_block && !wfused) MCMB_TRY(wstage(b));
        // stage 1
        if (one_block) MCMB_TRY(phase("l3 contraction (row moments)"));
        MCMB_TRY(launch_guard("row moments", (unsigned long long)T.ngroups, l3mfma ? 64 * nwv : 256));
#define SEP_ROWMOM(NN, GG, FF, HH)                                                                            \
    hipLaunchKernelGGL((sep_rowmom<NN, GG, FF, HH>), dim3((unsigned)T.ngroups), dim3(NN), 0, c->stream,      \
                       tab + T.o_groups, ginfo, wofs, (const double*)c->sg_wd.as<double>(), T.row0, L.kstep,  \
                       dU1, ldu, NB, ldB, c->sg_M.as<double>())
#define SEP_ROWMOM_MFMA(CC)                                                                                   \
    hipLaunchKernelGGL((sep_rowmom_mfma<CC, SR_GM / 16>), dim3((unsigned)T.ngroups), dim3(64 * nwv), 0, c->stream,       \
                       tab + T.o_groups, ginfo, wofs, (const double*)c->sg_wd.as<double>(), T.row0, L.kstep,  \
                       dU1, ldu, NB, ldB, c->sg_M.as<double>())
#define SEP_ROWMOM_LW(CC, AA, NN, SS)                                                                             \
    hipLaunchKernelGGL((sep_rowmom_lw<CC, AA, NN, SS>),                                                     \
                       dim3((unsigned)(8 * SR_XC * ((T.ngroups + 8 * SR_XC - 1) / (8 * SR_XC)))), dim3(64 * NN), \
                       0, c->stream, tab + T.o_groups, ginfo, (const RowDesc*)L.d_rows.as<RowDesc>(),          \
                       (const double*)L.d_ftp.as<double>() + ftab_guard(L.lmax), (const double*)L.d_vp.as<double>(), \
                       T.row0,                                                                                  \
                       T.ngroups, dU1, ldu, NB, ldB, c->sg_M.as<double>())
#define SEP_LW(CC, NN)                                                                                        \
    do {                                                                                                      \
        if (T.swap) {                                                                                         \
            if (wkind == 0) SEP_ROWMOM_LW(CC, 0, NN, true);                                                   \
            else SEP_ROWMOM_LW(CC, 1, NN, true);                                                              \
        } else {                                                                                              \
            if (wkind == 0) SEP_ROWMOM_LW(CC, 0, NN, false);                                                  \
            else SEP_ROWMOM_LW(CC, 1, NN, false);                                                             \
        }                                                                                                     \
    } while (0)
        if (wfused) {
            // the (column tiles, waves) pairs the width selection above produces
            if (nwv == 1) {
                if (cpw == 4) SEP_LW(4, 1);
                else if (cpw == 6) SEP_LW(6, 1);
                else SEP_LW(8, 1);
            } else if (nwv == 2) {
                if (cpw == 6) SEP_LW(6, 2);
                else SEP_LW(8, 2);
            } else {
                if (cpw == 6) SEP_LW(6, 4);
                else SEP_LW(8, 4);
            }
        } else if (l3mfma) {
            if (cpw == 4) SEP_ROWMOM_MFMA(4);
            else if (cpw == 6) SEP_ROWMOM_MFMA(6);
            else SEP_ROWMOM_MFMA(8);
        } else if (NB <= 256) SEP_ROWMOM(256, 1, 1, 1);
        else if (NB <= 512) SEP_ROWMOM(256, 1, 2, 1);
        else SEP_ROWMOM(256, 1, 3, 2);
#undef SEP_LW
#undef SEP_ROWMOM_LW
#undef SEP_ROWMOM_MFMA
#undef SEP_ROWMOM
        MCMB_HIP(hipGetLastError());
        // stages 2 and 3
#define SEP_GEMMF(CG, ST, FM, FN, NT, GSZ, AP, LDA, AIDX, BP, LDB, CP, LDC, MD, ND, TASK, TILE, ACC)           \
    hipLaunchKernelGGL((sep_gemm<CG, ST, FM, FN>), dim3(sep_blocks(NT, GSZ)), dim3(256), 0, c->stream,           \
                       (const double*)(AP), (int)(LDA), AIDX, (const double*)(BP), (long)(LDB), CP, (long)(LDC), MD, \
                       ND, reinterpret_cast<const SepTask*>(TASK), reinterpret_cast<const int4*>(TILE), NT, GSZ, ACC)
#define SEP_GEMM(CG, ST, ...)                                                                                 \
    do {                                                                                                      \
        if (small) SEP_GEMMF(CG, ST, 2, 2, __VA_ARGS__);                                                      \
        else SEP_GEMMF(CG, ST, SG_FM, SG_FN, __VA_ARGS__);                                                    \
    } while (0)
        if (one_block) MCMB_TRY(phase("l2 contraction (FP64 MFMA)"));
        MCMB_TRY(launch_guard("l2 contraction", sep_blocks(T.nt2, GS), 256));
        if (T.contig2)
            SEP_GEMM(true, 2, T.nt2, GS, dU, ldu, tab + T.o_aidx2, c->sg_M.as<double>(), ldB, c->sg_Y.as<double>(), ldB,
                     NB, NB, tab + T.o_task2, tab + T.o_tile2, 0);
        else
            SEP_GEMM(false, 2, T.nt2, GS, dU, ldu, tab + T.o_aidx2, c->sg_M.as<double>(), ldB, c->sg_Y.as<double>(),
                     ldB, NB, NB, tab + T.o_task2, tab + T.o_tile2, 0);
        MCMB_HIP(hipGetLastError());
        // Ys[i][{e,f}] = Y[i][e][f] + Y[i][f][e]
        if (one_block) MCMB_TRY(phase("symmetrisation"));
        {
            const int nb32 = (NB + 31) / 32;
            const unsigned nbp = (unsigned)(nb32 * (nb32 + 1) / 2);
            if (T.nruns > 65535) return fail(MCMB_ERR_ARG, "separable Gram: %d l1 runs in one block", T.nruns);
            MCMB_TRY(launch_guard("symmetrisation", (unsigned long long)nbp * T.nruns, 256));
            hipLaunchKernelGGL(sep_ysym, dim3(nbp, (unsigned)T.nruns), dim3(256), 0, c->stream,
                               (const double*)c->sg_Y.as<double>(), NPL, ldB, NB, nb32, c->sg_Ys.as<double>(), ldq2);
            MCMB_HIP(hipGetLastError());
        }
        if (one_block) MCMB_TRY(phase("l1 contraction (FP64 MFMA)"));
        MCMB_TRY(launch_guard("l1 contraction", sep_blocks(T.nt3, GS), 256));
        if (T.contig3)
            SEP_GEMM(true, 3, T.nt3, GS, dU, ldu, tab + T.o_aidx3, c->sg_Ys.as<double>(), ldq2, c->sg_G.as<double>(),
                     ldq2, NB, (int)NBP, tab + T.o_task3, tab + T.o_tile3, b > 0 ? 1 : 0);
        else
            SEP_GEMM(false, 3, T.nt3, GS, dU, ldu, tab + T.o_aidx3, c->sg_Ys.as<double>(), ldq2, c->sg_G.as<double>(),
                     ldq2, NB, (int)NBP, tab + T.o_task3, tab + T.o_tile3, b > 0 ? 1 : 0);
        MCMB_HIP(hipGetLastError());
        launches += 4;
    }
    if (comp) {
        // H over the compressed basis (r^3 entries), then back to pairs:
        // T1[(a,b)][z] = sum_g H[g][(a,b)] Ct[g][z] (into M), Y2s[a][{y,z}] =
        // sum_b Ct[b][y] T1[(a,b)][z] for y <= z (into Ys, pair order),
        // G[x][{y,z}] = sum_a Ct[a][x] Y2s[a][{y,z}] (into G): the fully
        // symmetrised tensor over pairs
        MCMB_TRY(phase("expansion to pairs"));
        const int r = NB;
        const int* tb = S.tab.as<int>();
        const long eh = (long)r * ldh;
        MCMB_TRY(launch_guard("expansion", (unsigned long long)((eh + 255) / 256), 256));
        hipLaunchKernelGGL(sep_hfull, dim3((unsigned)((eh + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)c->sg_G.as<double>(), ldq2, (long)NB * ldq2,
                           ksplit, r, S.h.as<double>(), ldh);
        MCMB_TRY(launch_guard("expansion", sep_blocks(S.nt[0], S.gs[0]), 256));
        SEP_GEMMF(true, 4, SG_FM, SG_FN, S.nt[0], S.gs[0], S.h.as<double>(), ldh, tb, S.ct.as<double>(), ld, c->sg_M.as<double>(),
                 ld, r * r, NP, tb + S.o_task[0], tb + S.o_tile[0], 0);
        MCMB_TRY(launch_guard("expansion", sep_blocks(S.nt[1], S.gs[1]), 256));
        hipLaunchKernelGGL((sep_gemm<true, 4, SG_FM, SG_FN, true>), dim3(sep_blocks(S.nt[1], S.gs[1])), dim3(256), 0,
                           c->stream, (const double*)S.ct.as<double>(), ld, tb, (const double*)c->sg_M.as<double>(),
                           (long)ld, c->sg_Ys.as<double>(), ldp, NP, NP,
                           reinterpret_cast<const SepTask*>(tb + S.o_task[1]),
                           reinterpret_cast<const int4*>(tb + S.o_tile[1]), S.nt[1], S.gs[1], 0);
        MCMB_TRY(launch_guard("expansion", sep_blocks(S.nt[2], S.gs[2]), 256));
        SEP_GEMMF(true, 4, SG_FM, SG_FN, S.nt[2], S.gs[2], S.ct.as<double>(), ld, tb, c->sg_Ys.as<double>(), ldp,
                 c->sg_G.as<double>(), ldp, NP, (int)NPP, tb + S.o_task[2], tb + S.o_tile[2], 0);
        MCMB_HIP(hipGetLastError());
        launches += 4;
    }
#undef SEP_GEMM
#undef SEP_GEMMF
    MCMB_TRY(phase("assembly"));
    if (comp) {
        // the 6-term sums straight from the expanded tensor, by 32 x 32 tiles
        // blocks of 6 x 6 tiles per XCD (assembly 0.205 -> 0.146 ms at 4960 modes;
        // 2 x 2: 0.165, 4 x 4: 0.152, 8 x 8: 0.147)
        constexpr int S = 6;
        const int nt = (n + 31) / 32, nbk = (nt + S - 1) / S;
        const unsigned long long ntile = ((unsigned long long)nbk * (nbk + 1) / 2 + 7) / 8 * 8 * S * S;
        MCMB_TRY(launch_guard("assembly", ntile, 256));
        hipLaunchKernelGGL(sep_assemble_sym<S>, dim3((unsigned)ntile), dim3(256), 0, c->stream,
                           (const double*)c->sg_G.as<double>(), ldp, R, L.d_modes.as<int>(), n, nt, d_gamma);
        MCMB_HIP(hipGetLastError());
        ++launches;
    } else {
        // Gs from G2, then the 6-term sums
        MCMB_TRY(launch_guard("symmetrised G", (unsigned long long)NPP, 256));
        hipLaunchKernelGGL(sep_gsym, dim3((unsigned)NPP), dim3(256), 0, c->stream, (const double*)c->sg_G.as<double>(),
                           ldp, NP, (const long*)c->sg_off.as<long>(), c->sg_Gs.as<double>());
        MCMB_HIP(hipGetLastError());
        const long ng = (long)n * n;
        MCMB_TRY(launch_guard("assembly", (unsigned long long)((ng + 255) / 256), 256));
        hipLaunchKernelGGL(sep_assemble, dim3((unsigned)((ng + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)c->sg_Gs.as<double>(), (const long*)c->sg_off.as<long>(), NP, R,
                           L.d_modes.as<int>(), n, d_gamma);
        MCMB_HIP(hipGetLastError());
        launches += 2;
    }
    if (np > 0) MCMB_HIP(hipEventRecord(c->evp[np], c->stream));
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    setup_done("launch to completion");
    setup_end(c);
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = launches;
    c->nphases = np;
    c->last_rank = comp ? NB : 0;
    return MCMB_OK;
}

}  // namespace mcmb
