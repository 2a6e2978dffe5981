// FP64 symmetric eigen-decomposition by Householder tridiagonalisation on
// the GPU (the eigen-decomposition of ortho_late_mpi.c:768-776 and
// ortho_prim_mpi.c:223-224, gsl_eigen_symmv + sort descending).
//
//   A = Q T Q^T (T tridiagonal, Q = H_0 ... H_{n-2}, H_j = I - tau_j v_j v_j^T)
//   T = Z diag(lambda) Z^T   (divide and conquer, dc_tridiag.hip; rocSOLVER
//                             dstedc when MCMB_TRIDIAG=rocsolver)
//   eigenvectors X = Q Z, applied as blocks of reflectors.
//
// Tridiagonalisation (the lower form of LAPACK dsytrd / dlatrd, in panels of
// TRD_NB columns).  A is kept full and symmetric, row-major; row j stands for
// column j.  Column j = k0 + i of a panel takes three kernels:
//   trd_row   x = A[j][j..n) - sum_{p<i} (V_p[c] W_p[j] + W_p[c] V_p[j])
//             (the panel's earlier reflectors, not yet in A), d_j = x_j, and
//             per workgroup its part of |x_{j+2..}|^2;
//   trd_symv  the reflector of x_{j+1..n), formed alike in every workgroup:
//             beta = -sign(x_{j+1}) |x|, tau = (beta - x_{j+1}) / beta,
//             v = x / (x_{j+1} - beta) with v_{j+1} = 1 (tau = 0 when
//             x_{j+2..n) = 0), e_j = beta; then y = A v over rows j+1.. and
//             per workgroup its parts of V_p^T v, W_p^T v and y^T v;
//   trd_symv_fin  those parts added: w = tau (y - V (W^T v) - W (V^T v)) +
//             alpha2 v with alpha2 = -tau/2 (w^T v), kept as y, the two dot
//             vectors and alpha2 until the next trd_row makes w final.
// Partial sums are added in workgroup order, so results do not depend on
// timing, and no workgroup waits for another.  At the end of a panel the
// trailing matrix takes A -= V W^T + W V^T on FP64 MFMA (lower tiles,
// mirrored, so A stays exactly symmetric).
//
// Back-transformation: per panel, H_k0 ... H_{k0+nb-1} = I - V T V^T (T upper
// triangular: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] V^T v_i, LAPACK
// dlarft), and with the eigenvectors as rows E (row k = z_k^T), X^T =
// Z^T H_{n-2} ... H_0, so panels from the last to the first apply
// E <- E - (E V) (T^T V^T), both products on FP64 MFMA.
#include "mcmb_internal.h"
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <dlfcn.h>
#include <algorithm>
#include <mutex>
#include <cstring>
#include <numeric>
#include <vector>

namespace mcmb {

constexpr int TRD_NB = 32;           // reflectors per panel
constexpr int TRD_MAXN = 8000;       // v of a column staged in LDS (under 64 KiB with the static part)
constexpr int TRD_ROWT = 256;        // trd_row threads
constexpr int TRD_SYMT = 512;        // trd_symv threads
constexpr int TRD_SYMG = 1024;       // most trd_symv workgroups

typedef double fd4 __attribute__((ext_vector_type(4)));

// Device-side state of the current column and panel.
struct TrdState {
    double tau;
    double alpha;               // x_{j+1}
    double c1[TRD_NB];          // V_p^T v
    double c2[TRD_NB];          // W_p^T v
    double alpha2[TRD_NB];      // w_p += alpha2[p] v_p, p of the panel
};

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// A = G (or D^-1 G D^-1 with norm[i] = sqrt G_ii).
__global__ void trd_init(const double* __restrict__ G, int n, int lda, int normalise, double* __restrict__ norm,
                         double* __restrict__ A)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long)n * n) {
        const int i = (int)(e / n), j = (int)(e % n);
        double v = G[e];
        if (normalise) v /= sqrt(G[(long)i * n + i]) * sqrt(G[(long)j * n + j]);
        A[(long)i * lda + j] = v;
    }
    if (norm && e < n) norm[e] = sqrt(G[e * n + e]);
}

// W_i[c] of the panel, final: tau_i (y_i[c] - sum_{p<i} (V_p[c] c2[p] +
// W_p[c] c1[p])) + alpha2_i V_i[c], from y_i = A v_i as trd_symv stored it in
// W_i and the final W_p (p < i).
__device__ __forceinline__ double trd_wfin(const double* __restrict__ Vp, const double* __restrict__ Wp, int n, int i,
                                          long c, double tau, const double* c1, const double* c2, double a2)
{
    double s = Wp[(long)i * n + c];
    for (int p = 0; p < i; ++p) s -= Vp[(long)p * n + c] * c2[p] + Wp[(long)p * n + c] * c1[p];
    return tau * s + a2 * Vp[(long)i * n + c];
}

// trd_symv's nb workgroup partials of column i's 1 + 2 i values (y^T v, then
// V_p^T v and W_p^T v) added in a fixed order by 256 threads: value q by the
// four threads 4 q .. 4 q + 3, each over the workgroups b = its index mod 4,
// then the four in a fixed pairing.  tot[q] (LDS) for q < 1 + 2 i; alpha2 =
// -tau/2 w^T v with w^T v = tau (y^T v - 2 c1 . c2).
__device__ __forceinline__ void trd_sum_parts(int i, int nb, const double* __restrict__ part, int tid, double* tot)
{
    const int q = tid >> 2, sub = tid & 3;
    double s = 0.0;
    if (q < 1 + 2 * i)
        for (int b = sub; b < nb; b += 4) s += part[(long)q * TRD_SYMG + b];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q < 1 + 2 * i && sub == 0) tot[q] = s;
}
__device__ __forceinline__ double trd_alpha2(int i, double tau, const double* tot)
{
    double cc = 0.0;
    for (int p = 0; p < i; ++p) cc = fma(tot[1 + p], tot[1 + i + p], cc);
    return -0.5 * tau * (tau * (tot[0] - 2.0 * cc));
}

// Column j = k0 + i.  W_{i-1} is made final on columns j+1.. (column j itself
// stays as trd_symv left it: nothing reads it again), then the row update
// x = A[j][j..) - sum_{p<i} (V_p W_p[j] + W_p V_p[j]), d_j, and each
// workgroup's part of |x_{j+2..}|^2 (trd_symv adds them and forms the
// reflector).
// Vp, Wp: the panel's rows (V_p at Vp + p n, W_p at Wp + p n).
__global__ __launch_bounds__(TRD_ROWT) void trd_row(const double* __restrict__ A, int lda, int n, int j, int i,
                                                   const double* __restrict__ Vp, double* __restrict__ Wp,
                                                   const double* __restrict__ tt, double* __restrict__ xbuf,
                                                   double* __restrict__ dd, TrdState* __restrict__ st,
                                                   const double* __restrict__ spart, int snb,
                                                   double* __restrict__ part)
{
    static_assert(TRD_ROWT == 256, "trd_sum_parts takes 256 threads");
    __shared__ double sV[TRD_NB], sW[TRD_NB];
    __shared__ double tot[2 * TRD_NB + 1];
    __shared__ double red[TRD_ROWT / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int q = i - 1;   // the column whose W becomes final here (none when i = 0)
    // this thread's column: its loads issued first, in flight during the sums below
    const int c = j + blockIdx.x * TRD_ROWT + tid;
    const bool cin = c < n;
    double x = cin ? A[(long)j * lda + c] : 0.0;
    double vp[TRD_NB - 1], wp[TRD_NB - 1];
#pragma unroll
    for (int p = 0; p < TRD_NB - 1; ++p) {
        vp[p] = (cin && p < q) ? Vp[(long)p * n + c] : 0.0;
        wp[p] = (cin && p < q) ? Wp[(long)p * n + c] : 0.0;
    }
    const double vq = (cin && q >= 0) ? Vp[(long)q * n + c] : 0.0;
    const double wqr = (cin && q >= 0) ? Wp[(long)q * n + c] : 0.0;
    // column q's sums, from trd_symv's partials (the same in every workgroup)
    if (q >= 0) trd_sum_parts(q, snb, spart, tid, tot);
    if (tid < i) sV[tid] = Vp[(long)tid * n + j];
    if (tid < q) sW[tid] = Wp[(long)tid * n + j];
    __syncthreads();
    const double* const c1 = tot + 1;
    const double* const c2 = tot + 1 + (q > 0 ? q : 0);
    double tq = 0.0, a2q = 0.0;
    if (q >= 0) {
        tq = tt[j - 1];
        a2q = trd_alpha2(q, tq, tot);
    }
    if (q >= 0 && w == 0) {
        // W_q[j] final: lane p < q adds its correction term, then the wave tree
        double t = 0.0;
        if (lane < q) t = Vp[(long)lane * n + j] * c2[lane] + sW[lane] * c1[lane];
        t = wave_sum(t);
        if (lane == 0) sW[q] = tq * (Wp[(long)q * n + j] - t) + a2q * sV[q];
    }
    __syncthreads();
    if (cin) {
        // one pass over the final W_p, V_p (p < q): the row update and W_q's correction
        double corr = 0.0;
#pragma unroll
        for (int p = 0; p < TRD_NB - 1; ++p) {
            if (p < q) {
                x -= vp[p] * sW[p] + wp[p] * sV[p];
                corr += vp[p] * c2[p] + wp[p] * c1[p];
            }
        }
        if (q >= 0) {
            const double wq = tq * (wqr - corr) + a2q * vq;
            if (c > j) Wp[(long)q * n + c] = wq;
            x -= vq * sW[q] + wq * sV[q];
        }
        if (c == j) dd[j] = x;
        if (c == j + 1) st->alpha = x;
        if (c >= j + 1) xbuf[c] = x;
    }
    const double xs = (cin && c >= j + 2) ? x : 0.0;
    {
        const double s = wave_sum(xs * xs);
        if (lane == 0) red[w] = s;
    }
    __syncthreads();
    if (tid == 0) {
        double s = 0.0;
        for (int k = 0; k < TRD_ROWT / 64; ++k) s += red[k];
        part[blockIdx.x] = s;
    }
}

// y = A v on rows j+1.. (the trailing matrix as of the panel's start; v in
// LDS; four rows per wave at a time, 16-byte loads kept four steps ahead in
// registers: lda is even, so the
// columns j+1+peel, j+2+peel of every row (peel = (j+1) & 1) are one aligned
// pair), stored in W_i, v stored as V_j; and the sums c1[p] = V_p^T v,
// c2[p] = W_p^T v (p < i, lane p of each wave) and y^T v over the
// workgroups.  The last workgroup forms alpha2 = -tau/2 w^T v with
// w^T v = tau (y^T v - 2 c1 . c2) (w = tau (y - V c2 - W c1) is made final
// by the next trd_row).
__global__ __launch_bounds__(TRD_SYMT) void trd_symv(const double* __restrict__ A, int lda, int n, int j, int i,
                                                    const double* __restrict__ xbuf, const double* __restrict__ Vp,
                                                    double* __restrict__ Wp, double* __restrict__ Vj,
                                                    const double* __restrict__ rpart, int nr, double* __restrict__ ee,
                                                    double* __restrict__ tt, TrdState* __restrict__ st,
                                                    double* __restrict__ part)
{
    constexpr int NWV = TRD_SYMT / 64, NQ = 2 * TRD_NB + 1, RW = 4, DP = 4;
    extern __shared__ double vs[];
    __shared__ double red[NWV][NQ];
    __shared__ double rs;
    const int m = n - j - 1, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    if (w == 0) {
        double s = 0.0;
        for (int b = lane; b < nr; b += 64) s += rpart[b];
        s = wave_sum(s);
        if (lane == 0) {
            const double alpha = st->alpha;
            double tau = 0.0, beta = alpha, scale = 0.0;
            if (s > 0.0) {
                beta = -copysign(sqrt(alpha * alpha + s), alpha);
                tau = (beta - alpha) / beta;
                scale = 1.0 / (alpha - beta);
            }
            rs = scale;
            if (blockIdx.x == 0) {
                ee[j] = beta;
                tt[j] = tau;
                st->tau = tau;
            }
        }
    }
    __syncthreads();
    const double scale = rs;
    for (int e = tid; e < m; e += TRD_SYMT) vs[e] = e == 0 ? 1.0 : scale * xbuf[j + 1 + e];
    __syncthreads();
    const int peel = (j + 1) & 1, npair = (m - peel) >> 1, tail = (m - peel) & 1;
    double a1 = 0.0, a2 = 0.0, yv = 0.0;   // lane p < i: V_p . v, W_p . v; lane 0: y . v
    const int nw = gridDim.x * NWV;
    for (int r0 = RW * (blockIdx.x * NWV + w); r0 < m; r0 += RW * nw) {
        const double2* Ap[RW];
        double y[RW];
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int rr = min(r0 + k, m - 1);
            const double* const Ar = A + (long)(j + 1 + rr) * lda + (j + 1);
            Ap[k] = reinterpret_cast<const double2*>(Ar + peel);
            y[k] = 0.0;
            if (lane == 0) {
                if (peel) y[k] = Ar[0] * vs[0];
                if (tail) y[k] = fma(Ar[m - 1], vs[m - 1], y[k]);
            }
        }
        // steps s of the wave (t = lane + 64 s; lanes past npair read the
        // last pair and weigh it by zero, so the loop is uniform), loaded DP
        // steps ahead into a ring of registers (slot d holds step s0 + d)
        const int ns = (npair + 63) >> 6;
        double2 ring[DP][RW];
        if (ns > 0) {
#pragma unroll
            for (int d = 0; d < DP; ++d)
#pragma unroll
                for (int k = 0; k < RW; ++k) ring[d][k] = Ap[k][min(lane + 64 * d, npair - 1)];
        }
        for (int s0 = 0; s0 < ns; s0 += DP) {
#pragma unroll
            for (int d = 0; d < DP; ++d) {
                const int st = s0 + d;
                if (st < ns) {
                    const int t = lane + 64 * st;
                    const bool ok = t < npair;
                    const int e = peel + 2 * min(t, npair - 1);
                    const double v0 = ok ? vs[e] : 0.0, v1 = ok ? vs[e + 1] : 0.0;
#pragma unroll
                    for (int k = 0; k < RW; ++k) {
                        y[k] = fma(ring[d][k].x, v0, y[k]);
                        y[k] = fma(ring[d][k].y, v1, y[k]);
                    }
                    if (st + DP < ns)
#pragma unroll
                        for (int k = 0; k < RW; ++k) ring[d][k] = Ap[k][min(t + 64 * DP, npair - 1)];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < RW; ++k) y[k] = wave_sum(y[k]);
#pragma unroll
        for (int k = 0; k < RW; ++k) {
            const int rr = r0 + k, r = j + 1 + rr;
            if (rr < m) {
                const double vr = vs[rr];
                if (lane < i) {
                    a1 = fma(Vp[(long)lane * n + r], vr, a1);
                    a2 = fma(Wp[(long)lane * n + r], vr, a2);
                }
                if (lane == 0) {
                    Wp[(long)i * n + r] = y[k];
                    Vj[r] = vr;
                    yv = fma(y[k], vr, yv);
                }
            }
        }
    }
    if (lane < i) {
        red[w][1 + lane] = a1;
        red[w][1 + i + lane] = a2;
    }
    if (lane == 0) red[w][0] = yv;
    __syncthreads();
    const int nq = 1 + 2 * i;
    // partial q of workgroup b at part[q TRD_SYMG + b]
    if (tid < nq) {
        double s = 0.0;
        for (int k = 0; k < NWV; ++k) s += red[k][tid];
        part[(long)tid * TRD_SYMG + blockIdx.x] = s;
    }
}

// The sums of a panel's last column (the next trd_row, which adds them itself
// for the other columns, belongs to the next panel): c1, c2 and alpha2 into
// the state for trd_wfinal.  tt: tau of the column.
__global__ __launch_bounds__(256) void trd_symv_fin(int i, int nb, const double* __restrict__ part,
                                                   TrdState* __restrict__ st, const double* __restrict__ tt)
{
    __shared__ double tot[2 * TRD_NB + 1];
    const int tid = threadIdx.x;
    trd_sum_parts(i, nb, part, tid, tot);
    __syncthreads();
    if (tid < i) {
        st->c1[tid] = tot[1 + tid];
        st->c2[tid] = tot[1 + i + tid];
    }
    if (tid == 0) st->alpha2[i] = trd_alpha2(i, tt[0], tot);
}

// W_q final (q = nb - 1, the panel's last column) on columns [c0, n).
__global__ void trd_wfinal(double* __restrict__ Wp, const double* __restrict__ Vp, int n, int q, int c0,
                           const double* __restrict__ tt, const TrdState* __restrict__ st)
{
    __shared__ double c1[TRD_NB], c2[TRD_NB];
    if (threadIdx.x < q) {
        c1[threadIdx.x] = st->c1[threadIdx.x];
        c2[threadIdx.x] = st->c2[threadIdx.x];
    }
    __syncthreads();
    const long c = c0 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    Wp[(long)q * n + c] = trd_wfin(Vp, Wp, n, q, c, tt[0], c1, c2, st->alpha2[q]);
}

// A[k1.., k1..] -= V W^T + W V^T (nb panel rows), 64 x 64 lower tiles I >= J
// on FP64 MFMA (four waves of 32 x 32), K = 2 nb; each tile subtracted from
// the lower block and, mirrored through LDS, from the upper one.
__global__ __launch_bounds__(256) void trd_syr2k(double* __restrict__ A, int lda, int n, int k1, const double* __restrict__ Vp,
                                                 const double* __restrict__ Wp, int nb)
{
    __shared__ double S[64][65];
    const int t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    const int J = t - I * (I + 1) / 2;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
    const int rl = 32 * (w >> 1), cl = 32 * (w & 1);                 // quadrant within the tile
    const int rb = k1 + 64 * I + rl, cb = k1 + 64 * J + cl;
    const int K = 2 * nb;
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (fd4){0.0, 0.0, 0.0, 0.0};
    for (int q0 = 0; q0 < K; q0 += 4) {
        const int q = q0 + kq;
        const bool qin = q < K;
        // A'(i, q) = V_q[i] | W_{q-nb}[i];  B'(j, q) = W_q[j] | V_{q-nb}[j]
        const double* const Ar = q < nb ? Vp + (long)q * n : Wp + (long)(q - nb) * n;
        const double* const Br = q < nb ? Wp + (long)q * n : Vp + (long)(q - nb) * n;
        double a[2], b[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int r = rb + 16 * x + l16, cc = cb + 16 * x + l16;
            a[x] = (qin && r < n) ? Ar[r] : 0.0;
            b[x] = (qin && cc < n) ? Br[cc] : 0.0;
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
    // element v of a 16 x 16 block: row 4 v + lane / 16, column lane % 16
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) S[rl + 16 * x + 4 * v + kq][cl + 16 * y + l16] = acc[x][y][v];
    __syncthreads();
    const int r0 = k1 + 64 * I, c0 = k1 + 64 * J;
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int ii = e >> 6, jj = e & 63;
        const int gi = r0 + ii, gj = c0 + jj;
        if (gi < n && gj < n && (I > J || jj <= ii)) A[(long)gi * lda + gj] -= S[ii][jj];
    }
    for (int e = threadIdx.x; e < 64 * 64; e += 256) {
        const int jj = e >> 6, ii = e & 63;
        const int gi = r0 + ii, gj = c0 + jj;
        if (gi < n && gj < n && (I > J || jj < ii)) A[(long)gj * lda + gi] -= S[ii][jj];
    }
}

__global__ void trd_identity(double* __restrict__ X, int n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long)n * n) X[e] = (e / n == e % n) ? 1.0 : 0.0;
}

__global__ void trd_lastd(const double* __restrict__ A, int lda, int n, double* __restrict__ dd)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) dd[n - 1] = A[(long)(n - 1) * lda + n - 1];
}

// Per panel b (one workgroup each): G = V V^T of its reflectors, T (dlarft,
// forward, column-wise) and Vt = T^T V (rows p: sum_{q <= p} T[q][p] v_q),
// over columns k0 + 1 .. n - 1 (the reflectors are zero before).
__global__ __launch_bounds__(1024) void trd_tmat(const double* __restrict__ V, const double* __restrict__ tt, int n,
                                                 double* __restrict__ Vt)
{
    __shared__ double Gs[TRD_NB][TRD_NB + 1], Ts[TRD_NB][TRD_NB + 1];
    const int k0 = blockIdx.x * TRD_NB, nb = min(TRD_NB, n - 1 - k0), c0 = k0 + 1;
    const int tid = threadIdx.x, a = tid / TRD_NB, b = tid % TRD_NB;
    if (a < nb && b < nb && b >= a) {
        double s = 0.0;
        const double* const va = V + (long)(k0 + a) * n;
        const double* const vb = V + (long)(k0 + b) * n;
        for (int c = k0 + b + 1; c < n; ++c) s = fma(va[c], vb[c], s);
        Gs[a][b] = s;
    }
    __syncthreads();
    if (tid < TRD_NB * TRD_NB) Ts[a][b] = 0.0;
    __syncthreads();
    if (tid < nb) {
        const int q = tid;
        Ts[q][q] = tt[k0 + q];
        for (int i = q + 1; i < nb; ++i) {
            double s = 0.0;
            for (int r = q; r < i; ++r) s = fma(Ts[q][r], Gs[r][i], s);   // G[r][i], r < i
            Ts[q][i] = -tt[k0 + i] * s;
        }
    }
    __syncthreads();
    for (int c = c0 + tid; c < n; c += 1024) {
        double s[TRD_NB];
#pragma unroll
        for (int p = 0; p < TRD_NB; ++p) s[p] = 0.0;
#pragma unroll
        for (int q = 0; q < TRD_NB; ++q) {
            const double vq = q < nb ? V[(long)(k0 + q) * n + c] : 0.0;
#pragma unroll
            for (int p = q; p < TRD_NB; ++p) s[p] = fma(Ts[q][p], vq, s[p]);
        }
#pragma unroll
        for (int p = 0; p < TRD_NB; ++p)
            if (p < nb) Vt[(long)(k0 + p) * n + c] = s[p];
    }
}

// M[r][p] = sum_{c >= c0} E[r][c] V_p[c] for 32 rows per workgroup and the
// panel's nb <= 32 reflectors (FP64 MFMA; the four waves split the columns c
// and add their 32 x 32 results in LDS in wave order).
__global__ __launch_bounds__(256) void trd_bt_proj(const double* __restrict__ E, int n, const double* __restrict__ Vp,
                                                   int nb, int c0, double* __restrict__ M)
{
    __shared__ double R[4][32][33];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
    const int r0 = blockIdx.x * 32;
    const int m = n - c0, seg = ((m + 3) / 4 + 3) / 4 * 4;   // columns per wave: a quarter, a multiple of 4
    const int ca = c0 + w * seg, cz = min(n, ca + seg);
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (fd4){0.0, 0.0, 0.0, 0.0};
    for (int k = ca; k < cz; k += 4) {
        const int c = k + kq;
        const bool cin = c < cz;
        double a[2], b[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int r = r0 + 16 * x + l16, p = 16 * x + l16;
            a[x] = (cin && r < n) ? E[(long)r * n + c] : 0.0;
            b[x] = (cin && p < nb) ? Vp[(long)p * n + c] : 0.0;
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) R[w][16 * x + 4 * v + kq][16 * y + l16] = acc[x][y][v];
    __syncthreads();
    for (int e = threadIdx.x; e < 32 * 32; e += 256) {
        const int rr = e >> 5, p = e & 31;
        const int r = r0 + rr;
        if (r < n && p < nb) M[(long)r * TRD_NB + p] = ((R[0][rr][p] + R[1][rr][p]) + R[2][rr][p]) + R[3][rr][p];
    }
}

// E[r][c] -= sum_p M[r][p] Vt_p[c] for c >= c0 (64 x 64 tiles, FP64 MFMA, K = nb).
__global__ __launch_bounds__(256) void trd_bt_update(double* __restrict__ E, int n, const double* __restrict__ M,
                                                     const double* __restrict__ Vt, int nb, int c0)
{
    const int m = n - c0, ntc = (m + 63) / 64;
    const int I = blockIdx.x / ntc, J = blockIdx.x % ntc;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
    const int rb = 64 * I + 32 * (w >> 1), cb = c0 + 64 * J + 32 * (w & 1);
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (fd4){0.0, 0.0, 0.0, 0.0};
    for (int p0 = 0; p0 < nb; p0 += 4) {
        const int p = p0 + kq;
        const bool pin = p < nb;
        double a[2], b[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int r = rb + 16 * x + l16, c = cb + 16 * x + l16;
            a[x] = (pin && r < n) ? M[(long)r * TRD_NB + p] : 0.0;
            b[x] = (pin && c < n) ? Vt[(long)p * n + c] : 0.0;
        }
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int r = rb + 16 * x + 4 * v + kq, c = cb + 16 * y + l16;
                if (r < n && c < n) E[(long)r * n + c] -= acc[x][y][v];
            }
}

// Rows of E in the order perm, each with its largest-magnitude component
// (the first on ties) made positive.
__global__ __launch_bounds__(256) void trd_sort_rows(const double* __restrict__ E, int n, const int* __restrict__ perm,
                                                     double* __restrict__ V)
{
    __shared__ double sv[256];
    __shared__ int si[256];
    const int k = blockIdx.x;
    const double* row = E + (long)perm[k] * n;
    double best = -1.0;
    int bi = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = fabs(row[i]);
        if (a > best) { best = a; bi = i; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            const double b = sv[threadIdx.x + s];
            const int j = si[threadIdx.x + s];
            if (b > sv[threadIdx.x] || (b == sv[threadIdx.x] && j < si[threadIdx.x])) {
                sv[threadIdx.x] = b;
                si[threadIdx.x] = j;
            }
        }
        __syncthreads();
    }
    const double sg = row[si[0]] < 0.0 ? -1.0 : 1.0;
    for (int i = threadIdx.x; i < n; i += 256) V[(long)k * n + i] = sg * row[i];
}

static int rb_status(rocblas_status s, const char* what)
{
    if (s == rocblas_status_success) return MCMB_OK;
    return fail(MCMB_ERR_HIP, "%s: rocBLAS/rocSOLVER status %d", what, (int)s);
}

// rocBLAS and rocSOLVER are loaded on the first eigen-decomposition (the
// solver library is large; processes that never call it do not map it).
struct RocSolver {
    bool ok = false;
    char err[256] = "";
    rocblas_status (*create_handle)(rocblas_handle*) = nullptr;
    rocblas_status (*destroy_handle)(rocblas_handle) = nullptr;
    rocblas_status (*set_stream)(rocblas_handle, hipStream_t) = nullptr;
    rocblas_status (*dstedc)(rocblas_handle, const rocblas_evect, const rocblas_int, double*, double*, double*,
                             const rocblas_int, rocblas_int*) = nullptr;
};

static const RocSolver& rocsolver_lib()
{
    static RocSolver R;
    static std::once_flag once;
    std::call_once(once, [] {
        void* hb = dlopen("librocblas.so.5", RTLD_NOW | RTLD_GLOBAL);
        void* hs = hb ? dlopen("librocsolver.so.0", RTLD_NOW | RTLD_GLOBAL) : nullptr;
        if (!hb || !hs) {
            snprintf(R.err, sizeof(R.err), "%s", dlerror());
            return;
        }
        R.create_handle = reinterpret_cast<decltype(R.create_handle)>(dlsym(hb, "rocblas_create_handle"));
        R.destroy_handle = reinterpret_cast<decltype(R.destroy_handle)>(dlsym(hb, "rocblas_destroy_handle"));
        R.set_stream = reinterpret_cast<decltype(R.set_stream)>(dlsym(hb, "rocblas_set_stream"));
        R.dstedc = reinterpret_cast<decltype(R.dstedc)>(dlsym(hs, "rocsolver_dstedc"));
        R.ok = R.create_handle && R.destroy_handle && R.set_stream && R.dstedc;
        if (!R.ok) snprintf(R.err, sizeof(R.err), "rocBLAS/rocSOLVER symbols missing");
    });
    return R;
}

// The reduction to tridiagonal form after trd_init: every column's and
// panel's launches on stream st_, then the last diagonal entry.
static int trd_reduce(TrdWork& W, int n, int lda, hipStream_t st_)
{
    DevBuf &A = W.A, &V = W.V, &Wp = W.Wp, &xb = W.xb, &dd = W.dd, &ee = W.ee, &tt = W.tt, &st = W.st,
           &part = W.part, &rpart = W.rpart;
    double* const Ad = A.as<double>();
    TrdState* const sd = reinterpret_cast<TrdState*>(st.p);
    int prev_g2 = 0;   // workgroups of the last trd_symv (its partials feed the next trd_row)
    for (int k0 = 0; k0 <= n - 2; k0 += TRD_NB) {
        const int nb = std::min(TRD_NB, n - 1 - k0), k1 = k0 + nb;
        const double* const Vp = V.as<double>() + (size_t)k0 * n;
        for (int i = 0; i < nb; ++i) {
            const int j = k0 + i, m = n - j - 1;
            const unsigned g1 = (unsigned)((n - j + TRD_ROWT - 1) / TRD_ROWT);
            hipLaunchKernelGGL(trd_row, dim3(g1), dim3(TRD_ROWT), 0, st_, (const double*)Ad, lda, n, j, i, Vp,
                               Wp.as<double>(), (const double*)tt.as<double>(), xb.as<double>(), dd.as<double>(), sd,
                               (const double*)part.as<double>(), prev_g2, rpart.as<double>());
            const unsigned g2 = (unsigned)std::min(TRD_SYMG, (m + 4 * (TRD_SYMT / 64) - 1) / (4 * (TRD_SYMT / 64)));
            hipLaunchKernelGGL(trd_symv, dim3(g2), dim3(TRD_SYMT), sizeof(double) * (size_t)m, st_,
                               (const double*)Ad, lda, n, j, i, (const double*)xb.as<double>(), Vp, Wp.as<double>(),
                               V.as<double>() + (size_t)j * n, (const double*)rpart.as<double>(), (int)g1,
                               ee.as<double>(), tt.as<double>(), sd, part.as<double>());
            prev_g2 = (int)g2;
            if (i == nb - 1)
                hipLaunchKernelGGL(trd_symv_fin, dim3(1), dim3(256), 0, st_, i, (int)g2,
                                   (const double*)part.as<double>(), sd, (const double*)(tt.as<double>() + j));
        }
        if (k1 < n) {
            hipLaunchKernelGGL(trd_wfinal, dim3((unsigned)((n - k1 + 255) / 256)), dim3(256), 0, st_,
                               Wp.as<double>(), Vp, n, nb - 1, k1, (const double*)(tt.as<double>() + k1 - 1),
                               (const TrdState*)sd);
            const int T = (n - k1 + 63) / 64;
            hipLaunchKernelGGL(trd_syr2k, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, st_, Ad, lda, n, k1, Vp,
                               (const double*)Wp.as<double>(), nb);
        }
        // alpha2 of the next panel starts from zero
        MCMB_HIP(hipMemsetAsync(sd->alpha2, 0, sizeof(double) * TRD_NB, st_));
    }
    hipLaunchKernelGGL(trd_lastd, dim3(1), dim3(64), 0, st_, (const double*)Ad, lda, n, dd.as<double>());
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

bool tridiag_supported(int n) { return n >= 1 && n <= TRD_MAXN; }

int tridiag_eigen(mcmb_ctx* c, int n, const double* d_G, int normalise, double* d_norm, double* evals, DevBuf& d_V)
{
    if (!tridiag_supported(n)) return fail(MCMB_ERR_ARG, "tridiagonal eigensolver: n = %d outside 1..%d", n, TRD_MAXN);
    const size_t nn = (size_t)n * n;
    const int lda = n + (n & 1);   // even: 16-byte aligned column pairs in every row
    const int npan = std::max(1, (n - 1 + TRD_NB - 1) / TRD_NB);
    TrdWork& W = c->trd;
    DevBuf &A = W.A, &V = W.V, &Wp = W.Wp, &xb = W.xb, &dd = W.dd, &ee = W.ee, &tt = W.tt, &st = W.st,
           &part = W.part, &rpart = W.rpart, &E = W.E, &Vt = W.Vt, &M = W.M, &perm = W.perm, &info = W.info;
    MCMB_TRY(A.alloc(sizeof(double) * (size_t)n * lda));
    MCMB_TRY(V.alloc(sizeof(double) * nn));
    MCMB_TRY(Wp.alloc(sizeof(double) * TRD_NB * (size_t)n));
    MCMB_TRY(xb.alloc(sizeof(double) * n));
    MCMB_TRY(dd.alloc(sizeof(double) * n));
    MCMB_TRY(ee.alloc(sizeof(double) * n));
    MCMB_TRY(tt.alloc(sizeof(double) * n));
    MCMB_TRY(st.alloc(sizeof(TrdState)));
    MCMB_TRY(part.alloc(sizeof(double) * (2 * TRD_NB + 1) * TRD_SYMG));
    MCMB_TRY(rpart.alloc(sizeof(double) * ((n + TRD_ROWT - 1) / TRD_ROWT)));
    MCMB_TRY(info.alloc(sizeof(int)));
    MCMB_HIP(hipMemsetAsync(V.p, 0, sizeof(double) * nn, c->stream));
    MCMB_HIP(hipMemsetAsync(Wp.p, 0, sizeof(double) * TRD_NB * (size_t)n, c->stream));
    MCMB_HIP(hipMemsetAsync(ee.p, 0, sizeof(double) * n, c->stream));
    MCMB_HIP(hipMemsetAsync(tt.p, 0, sizeof(double) * n, c->stream));
    MCMB_HIP(hipMemsetAsync(st.p, 0, sizeof(TrdState), c->stream));
    MCMB_TRY(launch_guard("tridiagonal init", (nn + 255) / 256, 256));
    hipLaunchKernelGGL(trd_init, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, d_G, n, lda, normalise,
                       d_norm, A.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_TRY(trd_reduce(W, n, lda, c->stream));
    // eigenpairs of T: eigenvalues ascending in dd, eigenvectors as the
    // columns of a column-major n x n matrix, i.e. the rows of E
    MCMB_TRY(E.alloc(sizeof(double) * nn));
    std::vector<double> w(n);
    const char* solver = getenv("MCMB_TRIDIAG");
    if (solver && !strcmp(solver, "rocsolver")) {
        const RocSolver& RS = rocsolver_lib();
        if (!RS.ok) return fail(MCMB_ERR_HIP, "tridiagonal eigensolver: cannot load rocBLAS/rocSOLVER (%s)", RS.err);
        if (!c->rb_handle) {
            rocblas_handle h = nullptr;
            MCMB_TRY(rb_status(RS.create_handle(&h), "rocblas_create_handle"));
            c->rb_handle = h;
        }
        rocblas_handle h = static_cast<rocblas_handle>(c->rb_handle);
        MCMB_TRY(rb_status(RS.set_stream(h, c->stream), "rocblas_set_stream"));
        hipLaunchKernelGGL(trd_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, E.as<double>(), n);
        MCMB_HIP(hipGetLastError());
        MCMB_TRY(rb_status(RS.dstedc(h, rocblas_evect_tridiagonal, n, dd.as<double>(), ee.as<double>(),
                                      E.as<double>(), n, info.as<int>()),
                           "rocsolver_dstedc"));
        int inf = 0;
        MCMB_HIP(hipMemcpyAsync(&inf, info.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipMemcpyAsync(w.data(), dd.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        if (inf != 0) return fail(MCMB_ERR_ARG, "tridiagonal eigensolver (dstedc) failed: info = %d at n = %d", inf, n);
    } else if (solver && strcmp(solver, "dc")) {
        return fail(MCMB_ERR_ARG, "MCMB_TRIDIAG=%s: expected dc or rocsolver", solver);
    } else {
        std::vector<double> dh(n), eh(n);
        MCMB_HIP(hipMemcpyAsync(dh.data(), dd.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipMemcpyAsync(eh.data(), ee.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        MCMB_TRY(dc_tridiag(c, n, dh.data(), eh.data(), E.as<double>(), w.data()));
    }
    // back-transformation, panels from the last to the first
    if (n >= 2) {
        MCMB_TRY(Vt.alloc(sizeof(double) * nn));
        MCMB_TRY(M.alloc(sizeof(double) * TRD_NB * (size_t)n));
        hipLaunchKernelGGL(trd_tmat, dim3((unsigned)npan), dim3(1024), 0, c->stream, (const double*)V.as<double>(),
                           (const double*)tt.as<double>(), n, Vt.as<double>());
        for (int b = npan - 1; b >= 0; --b) {
            const int k0 = b * TRD_NB, nb = std::min(TRD_NB, n - 1 - k0), c0 = k0 + 1;
            hipLaunchKernelGGL(trd_bt_proj, dim3((unsigned)((n + 31) / 32)), dim3(256), 0, c->stream,
                               (const double*)E.as<double>(), n, (const double*)(V.as<double>() + (size_t)k0 * n), nb,
                               c0, M.as<double>());
            const unsigned tiles = (unsigned)(((n + 63) / 64) * ((n - c0 + 63) / 64));
            hipLaunchKernelGGL(trd_bt_update, dim3(tiles), dim3(256), 0, c->stream, E.as<double>(), n,
                               (const double*)M.as<double>(), (const double*)(Vt.as<double>() + (size_t)k0 * n), nb, c0);
        }
        MCMB_HIP(hipGetLastError());
    }
    // descending order
    std::vector<int> pm(n);
    std::iota(pm.begin(), pm.end(), 0);
    std::stable_sort(pm.begin(), pm.end(), [&](int a, int b) { return w[a] > w[b]; });
    for (int k = 0; k < n; ++k) evals[k] = w[pm[k]];
    MCMB_TRY(perm.alloc(sizeof(int) * n));
    MCMB_HIP(hipMemcpyAsync(perm.p, pm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(d_V.alloc(sizeof(double) * nn));
    hipLaunchKernelGGL(trd_sort_rows, dim3(n), dim3(256), 0, c->stream, (const double*)E.as<double>(), n,
                       perm.as<int>(), d_V.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));   // pm and the work buffers go out of scope
    return MCMB_OK;
}

void tridiag_release(mcmb_ctx* c)
{
    // a handle exists only if the library was loaded
    if (c->rb_handle) (void)rocsolver_lib().destroy_handle(static_cast<rocblas_handle>(c->rb_handle));
    c->rb_handle = nullptr;
    c->trd = TrdWork();
}

}  // namespace mcmb
