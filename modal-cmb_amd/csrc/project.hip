// Shape projection on the Gauss-Legendre k-tetrapyd, with tangent (forward)
// and adjoint (reverse) derivatives in (omega, phi), for the feature shape
// S = sin(omega (k1+k2+k3) + phi) / (k1 k2 k3)^2 (calculate_shape.c:933-946).
//
//   alpha_n(theta) = sum_{cube} W_ijk sin(omega K_ijk + phi) Q_n(i,j,k)
//   W_ijk = mask_ijk w_i w_j w_k / (k_i + k_j + k_k)
// restates calculate_eigen (calculate_eigen.c:86-145) on the GL grid: its
// integrand is shape3/K = (k1k2k3)^2 S / K.  W and S are symmetric, so
// alpha_n = A[r][s][t] with A_rst = sum_ijk F_ijk q_r(i) q_s(j) q_t(k), F = W S.
//
// With A_ij = omega (k_i + k_j) + phi and B_k = omega k_k,
//   sin(A + B) = sin A cos B + cos A sin B,
// the k-sum separates:  Y[(i,j)][c,t] = sum_k W_ijk Z[k][c,t] with the four
// column sets Z = {cos B, sin B, k cos B, k sin B} q_t(k)  -- one MFMA GEMM
// (W resident in HBM, N^2 x N) for a batch of shapes.  Then per shape
//   G0 = sA Yc + cA Ys                         (value,  sin(A+B))
//   G2 = cA Yc - sA Ys                         (d/dphi, cos(A+B))
//   G1 = (k_i + k_j) G2 + cA Yc' - sA Ys'      (d/domega, K cos(A+B))
//   H_x[i][s][t] = sum_j q_s(j) G_x[i][j][t],  alpha_x[n] = sum_i q_r(i) H_x[i][s][t].
// The adjoint (VJP) takes gbar = dL/dalpha, forms Hbar[i][s][t] =
// sum_r q_r(i) Abar[r][s][t] (Abar = gbar scattered onto the canonical
// triples), Gbar[i][j][t] = sum_s q_s(j) Hbar[i][s][t], and reduces
// dL/dphi = sum Gbar G2, dL/domega = sum Gbar G1.
#include "mcmb_internal.h"
#include <type_traits>
#include <cstring>
#include <algorithm>
#include <cmath>
#include <vector>

namespace mcmb {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int PT = 128;     // GEMM tile (M and N)
constexpr int PK = 16;      // GEMM K chunk
constexpr int PBATCH = 32;  // shapes per batch

// ---- full-cube weight ------------------------------------------------------
// W[(i*N + j)*Kp + k]; the mask compares the sorted triple exactly as the
// oracle does: k_c <= k_a + k_b (FP64).
// Row of the pair (i, j) in the packed i <= j order of the stage-1 product.
__device__ __forceinline__ long pair_row(int i, int j, int N)
{
    const int lo = min(i, j), hi = max(i, j);
    return (long)lo * N - (long)lo * (lo - 1) / 2 + (hi - lo);
}

template <class T>
__global__ void proj_build_W(const double* __restrict__ kn, const double* __restrict__ kw, int N, int Kp,
                             const int2* __restrict__ rowij, long M, T* __restrict__ W)
{
    // W[(i,j) pair row][k], pair rows i <= j only (W is symmetric in i, j)
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t tot = (size_t)M * Kp;
    if (e >= tot) return;
    int k = (int)(e % Kp);
    const int2 ij = rowij[e / Kp];
    double w = 0.0;
    if (k < N) {
        int a = ij.x, b = ij.y, c = k, t;
        if (a > b) { t = a; a = b; b = t; }
        if (b > c) { t = b; b = c; c = t; if (a > b) { t = a; a = b; b = t; } }
        if (kn[c] <= kn[a] + kn[b]) w = kw[a] * kw[b] * kw[c] / (kn[a] + kn[b] + kn[c]);
    }
    W[e] = (T)w;
}

// The K chunks of each PT-row tile of W that hold a nonzero weight: tiles[t]
// = (first chunk, chunk count) over the tile's rows (an all-zero tile keeps
// chunk 0).  The stage-1 GEMMs run only over those chunks: the others add
// exact zeros.
template <class T>
__global__ __launch_bounds__(256) void proj_w_tiles(const T* __restrict__ W, long M, int Kp,
                                                    int2* __restrict__ tiles)
{
    __shared__ int lo, hi;
    if (threadIdx.x == 0) {
        lo = Kp;
        hi = -1;
    }
    __syncthreads();
    const long r0 = (long)blockIdx.x * PT, ne = (long)min((long)PT, M - r0) * Kp;
    int mlo = Kp, mhi = -1;
    for (long e = threadIdx.x; e < ne; e += 256)
        if (W[r0 * Kp + e] != (T)0) {
            const int k = (int)(e % Kp);
            mlo = min(mlo, k);
            mhi = max(mhi, k);
        }
    atomicMin(&lo, mlo);
    atomicMax(&hi, mhi);
    __syncthreads();
    if (threadIdx.x == 0) {
        const int c0 = hi < 0 ? 0 : lo / PK, c1 = hi < 0 ? 0 : hi / PK;
        tiles[blockIdx.x] = make_int2(c0, c1 - c0 + 1);
    }
}

// Z[k][col], col = (b*2 + c)*R + t for shapes b < nb of the batch: c = 0
// cos(omega_b k) q_t(k), c = 1 sin(omega_b k) q_t(k); Kp x ldz.
template <class T>
__global__ void proj_build_Z(const double* __restrict__ kn, const double* __restrict__ q, int ldq, int N, int R,
                             const double* __restrict__ theta, int nb, int Kp, int ldz, T* __restrict__ Z)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)Kp * ldz) return;
    int k = (int)(e / ldz), col = (int)(e % ldz);
    double z = 0.0;
    if (k < N && col < nb * 2 * R) {
        int t = col % R, c = (col / R) % 2, b = col / (2 * R);
        double om = theta[2 * b];
        double sb, cb;
        sincos(om * kn[k], &sb, &cb);
        z = (c ? sb : cb) * q[(size_t)k * ldq + t];
    }
    Z[e] = (T)z;
}

// ---- MFMA GEMM  C[M][Nc] = A[M][Kp] B[Kp][ldb]  (A row-major lda = Kp) ------
template <class T> struct MF;
template <> struct MF<double> {
    typedef d4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c)
    { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    // output row of element v for lane: 4v + lane/16
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * v + (lane >> 4); }
};
template <> struct MF<float> {
    typedef f4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c)
    { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    // output row of element v for lane: 4 (lane/16) + v
    static __device__ __forceinline__ int row(int v, int lane) { return 4 * (lane >> 4) + v; }
};

template <class T>
__global__ __launch_bounds__(256, 2) void proj_gemm(const T* __restrict__ A, const T* __restrict__ B, int M, int Kp,
                                                    int Nc, int ldb, T* __restrict__ C, int ldc)
{
    __shared__ T As[2][PK][PT + 4];
    __shared__ T Bs[2][PK][PT + 4];
    const int m0 = blockIdx.y * PT, n0 = blockIdx.x * PT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    typename MF<T>::V acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[a][b][v] = 0;
    // staging: thread loads 8 elements of A (rows r, k chunk) and 8 of B
    T ra[8], rb[8];
    auto gload = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int e = tid + u * 256;                  // 0..2047
            int r = e >> 4, kk = e & 15;            // A: 128 rows x 16
            int gm = m0 + r;
            ra[u] = (gm < M) ? A[(size_t)gm * Kp + k0 + kk] : (T)0;
            int kb = e >> 7, cc = e & 127;          // B: 16 x 128
            int gn = n0 + cc;
            rb[u] = (gn < Nc) ? B[(size_t)(k0 + kb) * ldb + gn] : (T)0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            int e = tid + u * 256;
            As[buf][e & 15][e >> 4] = ra[u];
            Bs[buf][e >> 7][e & 127] = rb[u];
        }
    };
    gload(0);
    sstore(0);
    __syncthreads();
    const int nk = Kp / PK;
    for (int kc = 0; kc < nk; ++kc) {
        const int buf = kc & 1;
        if (kc + 1 < nk) gload((kc + 1) * PK);
#pragma unroll
        for (int ks = 0; ks < PK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            T av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][kk][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[buf][kk][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = MF<T>::mma(av[a], bv[b], acc[a][b]);
        }
        if (kc + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int gm = m0 + wm * 64 + a * 16 + MF<T>::row(v, lane);
                int gn = n0 + wn * 64 + b * 16 + (lane & 15);
                if (gm < M && gn < Nc) C[(size_t)gm * ldc + gn] = acc[a][b][v];
            }
}

// sin and cos of omega_b k_j for the shapes of a batch: scb[b][j] = (sin, cos).
// Stage 2 and the adjoint form sin / cos of A_ij = (omega k_i + phi) +
// omega k_j by the angle-addition formulas from these and one sincos per
// workgroup.
__global__ void proj_sincos_tab(const double* __restrict__ kn, const double* __restrict__ theta, int N, int nb,
                                double* __restrict__ scb)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * N) return;
    const int b = e / N, j = e % N;
    double sb, cb;
    sincos(theta[2 * b] * kn[j], &sb, &cb);
    scb[2 * e] = sb;
    scb[2 * e + 1] = cb;
}

// ---- forward stage 2: per (shape b, i) -------------------------------------
// H[b][x][i][s][t] = sum_j q_s(j) G_x(i, j, t) with G_x(i, j, t) = sum_k
// W_ijk {sin, cos}(omega K + phi) q_t(k), x = 0 (sin: the value) and x = 1
// (cos: the derivatives).  From stage 1's Yc, Ys (cos / sin omega k_k
// columns) by angle addition: G_0 = sin A Yc + cos A Ys, G_1 = cos A Yc -
// sin A Ys, A = omega (k_i + k_j) + phi.  Y's row of the pair (i, j) is
// rowof[pair_row(i, j)] (stage 1 runs over W's rows in their own order).
template <class T>
__global__ __launch_bounds__(256) void proj_stage2(const T* __restrict__ Y, int ldy, const int* __restrict__ rowof,
                                                   const double* __restrict__ kn,
                                                   const double* __restrict__ q, int ldq, int N, int R,
                                                   const double* __restrict__ theta, int nb,
                                                   const double* __restrict__ scb, double* __restrict__ H)
{
    // one workgroup per (shape b, node i); H_x[s][t] = sum_j q_s(j) G_x[j][t]
    // over chunks of 32 nodes j, on FP64 MFMA 16x16x4: A = Q^T (s x j), B =
    // G_x (j x t).  Wave w owns the s tile w / 2 and the t tile w % 2 of both
    // x (2 accumulators, the A operand shared).  The thread's 4 elements (j =
    // j0 + tid / 32 + 8 u, t = tid % 32) of the next chunk are loaded (Y, q,
    // sin / cos omega k_j) while the current chunk's products run.
    __shared__ double G[2][32][32];     // [x][j in chunk][t]
    __shared__ double Qs[32][32];       // [j in chunk][s]
    const int b = blockIdx.x % nb, i = blockIdx.x / nb;
    const int tid = threadIdx.x, t = tid & 31, jr = tid >> 5;
    const int lane = tid & 63, w = tid >> 6, sm = 16 * (w >> 1), tn = 16 * (w & 1);
    const double ki = kn[i];
    double sa, ca;   // sin, cos of omega k_i + phi
    sincos(theta[2 * b] * ki + theta[2 * b + 1], &sa, &ca);
    const double* const sc = scb + 2 * (size_t)b * N;
    d4 acc[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) acc[x] = (d4){0.0, 0.0, 0.0, 0.0};
    T yv[4][2];
    double qv[4], sv[4], cv[4];
    // loads of chunk j0 (nodes past N repeat the last one and are zeroed in stage())
    auto fetch = [&](int j0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int j = min(j0 + jr + 8 * u, N - 1), tt = min(t, R - 1);
            const T* y = Y + (size_t)rowof[pair_row(i, j, N)] * ldy + (size_t)b * 2 * R + tt;
            yv[u][0] = y[0];
            yv[u][1] = y[R];
            qv[u] = q[(size_t)j * ldq + tt];
            sv[u] = sc[2 * j];
            cv[u] = sc[2 * j + 1];
        }
    };
    auto stage = [&](int j0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int jj = jr + 8 * u;
            const bool in = j0 + jj < N && t < R;
            const double sA = sa * cv[u] + ca * sv[u], cA = ca * cv[u] - sa * sv[u];
            const double yc = (double)yv[u][0], ys = (double)yv[u][1];
            const double g0 = sA * yc + cA * ys;
            const double g1 = cA * yc - sA * ys;
            G[0][jj][t] = in ? g0 : 0.0;
            G[1][jj][t] = in ? g1 : 0.0;
            Qs[jj][t] = in ? qv[u] : 0.0;
        }
    };
    fetch(0);
    for (int j0 = 0; j0 < N; j0 += 32) {
        __syncthreads();   // the previous chunk's products are done with G and Qs
        stage(j0);
        __syncthreads();
        fetch(min(j0 + 32, N - 1));
        // 8 K-steps of 4 nodes (those past N have G = Q = 0)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int jj = 4 * ks + (lane >> 4);
            const double a = Qs[jj][sm + (lane & 15)];
#pragma unroll
            for (int x = 0; x < 2; ++x)
                acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, G[x][jj][tn + (lane & 15)], acc[x], 0, 0, 0);
        }
    }
    // element v of a lane: s = sm + 4 v + lane / 16, t = tn + lane % 16
    const int tt = tn + (lane & 15);
    if (tt < R)
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int ss = sm + 4 * v + (lane >> 4);
                if (ss < R) H[(((size_t)b * 2 + x) * N + i) * R * R + ss * R + tt] = acc[x][v];
            }
}

// D[b][x][r][st] (st = s R + t) on FP64 MFMA 16x16x4:
//   x = 0: sum_i q_r(i) H[b][0][i][st]       (alpha)
//   x = 1: sum_i q_r(i) H[b][1][i][st]       (d alpha / d phi)
//   x = 2: sum_i k_i q_r(i) H[b][1][i][st]   (E, the k_i-weighted cos form)
// One wave per (shape b, H block y, block of 16 columns st): y = 0 makes x =
// 0, y = 1 makes x = 1 and 2 from the same B loads (each H element is read
// once); the two r tiles share them too.  K runs over the nodes i, the loads
// of 4 K-steps issued before their products.
//
// E gives d alpha / d omega = sum W K cos(omega K + phi) q_r q_s q_t: W and
// the cos term are symmetric in the three legs, so each of k_j and k_k can
// be moved onto leg i by relabelling, and d alpha / d omega [r][s][t] =
// E[r][s][t] + E[s][r][t] + E[t][s][r].
__global__ __launch_bounds__(64) void proj_stage3(const double* __restrict__ H, const double* __restrict__ q, int ldq,
                                                  const double* __restrict__ kn, int N, int R, double* __restrict__ D)
{
    const int lane = threadIdx.x;
    const int RR = R * R, nblk = (RR + 15) / 16;
    const int by = blockIdx.x / nblk, st = 16 * (blockIdx.x % nblk) + (lane & 15);
    const int b = by >> 1, y = by & 1;   // y uniform over the wave
    const double* const h = H + ((size_t)b * 2 + y) * N * RR + min(st, RR - 1);
    const int r0 = lane & 15, r1 = 16 + (lane & 15);
    d4 acc[2][2];   // [w: 0 plain, 1 k-weighted][r tile]
#pragma unroll
    for (int w = 0; w < 2; ++w)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) acc[w][rt] = (d4){0.0, 0.0, 0.0, 0.0};
    const int nks = (N + 3) / 4;
    constexpr int U = 4;
    for (int k0 = 0; k0 < nks; k0 += U) {
        double bv[U], a0[U], a1[U], kf[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            // K-steps past the last are zero through bv (i >= N)
            const int i = 4 * (k0 + u) + (lane >> 4);
            const int ic = min(i, N - 1);
            bv[u] = i < N ? h[(size_t)ic * RR] : 0.0;
            a0[u] = r0 < R ? q[(size_t)ic * ldq + r0] : 0.0;
            a1[u] = r1 < R ? q[(size_t)ic * ldq + r1] : 0.0;
            kf[u] = kn[ic];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], bv[u], acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], bv[u], acc[0][1], 0, 0, 0);
            if (y) {
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(kf[u] * a0[u], bv[u], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(kf[u] * a1[u], bv[u], acc[1][1], 0, 0, 0);
            }
        }
    }
    // element v of tile rt: r = 16 rt + 4 v + lane / 16, column st; x = y + w
    if (st < RR)
#pragma unroll
        for (int w = 0; w < 2; ++w)
            if (w <= y)
#pragma unroll
                for (int rt = 0; rt < 2; ++rt)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int r = 16 * rt + 4 * v + (lane >> 4);
                        if (r < R) D[(((size_t)b * 3 + y + w) * R + r) * RR + st] = acc[w][rt][v];
                    }
}

// alpha[b][m] = D[b][0][r][s R + t], d alpha / d omega = E[r][s][t] +
// E[s][r][t] + E[t][s][r] (E = D[b][2]), d alpha / d phi = D[b][1][r][s R +
// t], (r, s, t) = order[m], written in the caller's layout: alpha[b][m] and
// dalpha[b][x][m] (x = 0: omega, 1: phi).
__global__ void proj_stage3_modes(const double* __restrict__ D, int R, const int* __restrict__ modes, int n, int nb,
                                  double* __restrict__ alpha, double* __restrict__ dalpha)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nb * 3 * n) return;
    const int m = e % n, x = (e / n) % 3, b = e / (3 * n);
    const int pk = modes[m], r = pk & 255, s = (pk >> 8) & 255, t = (pk >> 16) & 255;
    const int RR = R * R;
    const double* const Db = D + (size_t)b * 3 * R * RR;
    if (x == 0) alpha[(size_t)b * n + m] = Db[(size_t)r * RR + s * R + t];
    else if (x == 1) dalpha[((size_t)b * 2 + 1) * n + m] = Db[(size_t)R * RR + (size_t)r * RR + s * R + t];
    else {
        const double* const E = Db + 2 * (size_t)R * RR;
        dalpha[(size_t)b * 2 * n + m] = E[(size_t)r * RR + s * R + t] + E[(size_t)s * RR + r * R + t] +
                                        E[(size_t)t * RR + s * R + r];
    }
}

// ---- adjoint ---------------------------------------------------------------
// With Gb[r][s][t] = gbar[b][m] at (r, s, t) = order[m] (0 elsewhere) and
// Gh[r][s][t] = Gb[r][s][t] + Gb[s][r][t] + Gb[t][s][r]:
//   dL/domega = sum Gb (E[r][s][t] + E[s][r][t] + E[t][s][r]) = sum Gh E,
//   dL/dphi   = sum Gb Dphi = sum Gh Dphi / 3
// (Dphi is symmetric in r, s, t, and Gh / 3 has the symmetric part of Gb).
// Both are then sum_i {k_i, 1/3} P_i with P_i = sum_{j,t} Gbar(i,j,t)
// G_1(i,j,t), Gbar(i,j,t) = sum_s q_s(j) Hbar[i][s][t], Hbar[i][s][t] =
// sum_r q_r(i) Gh[r][s][t].
//
// Hbar[b][i][s][t]: per shape a product Q (N x R) . Gh (R x R^2) on FP64
// MFMA 16x16x4, Gh gathered through slot[r][s][t] = m (-1: no mode).  One
// wave per (shape b, block of 16 columns st): its B operand (Gh, R padded to
// 32) is gathered once into registers, then the N / 16 row tiles of Q run
// through 8 K-steps each.
__global__ __launch_bounds__(64) void proj_hbar(const double* __restrict__ gbar, const int* __restrict__ slot,
                                                const double* __restrict__ q, int ldq, int N, int R, int n, int nb,
                                                double* __restrict__ Hb)
{
    const int lane = threadIdx.x;
    const int RR = R * R, nblk = (RR + 15) / 16;
    const int b = blockIdx.x / nblk, st = 16 * (blockIdx.x % nblk) + (lane & 15);
    double bv[8];
    const int s = min(st, RR - 1) / R, t = min(st, RR - 1) % R;
    auto g = [&](int m) { return m >= 0 ? gbar[(size_t)b * n + m] : 0.0; };
#pragma unroll
    for (int ks = 0; ks < 8; ++ks) {
        const int r = 4 * ks + (lane >> 4);
        bv[ks] = 0.0;
        if (r < R && st < RR)
            bv[ks] = g(slot[r * RR + s * R + t]) + g(slot[s * RR + r * R + t]) + g(slot[t * RR + s * R + r]);
    }
    for (int i0 = 0; i0 < N; i0 += 16) {
        const int ia = min(i0 + (lane & 15), N - 1);   // A row (rows past N: the last, not stored)
        d4 acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int r = 4 * ks + (lane >> 4);
            const double a = r < R ? q[(size_t)ia * ldq + r] : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv[ks], acc, 0, 0, 0);
        }
        // element v: row i0 + 4 v + lane / 16, column st
        if (st < RR)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = i0 + 4 * v + (lane >> 4);
                if (i < N) Hb[((size_t)b * N + i) * RR + st] = acc[v];
            }
    }
}

// dpart[b][i] = (k_i P_i, P_i / 3), P_i = sum_{j,t} Gbar(i,j,t) G_1(i,j,t), for
// one (b, i) per workgroup
template <class T>
__global__ __launch_bounds__(256) void proj_adjoint(const T* __restrict__ Y, int ldy, const int* __restrict__ rowof,
                                                    const double* __restrict__ kn,
                                                    const double* __restrict__ q, int ldq, int N, int R,
                                                    const double* __restrict__ theta, int nb,
                                                    const double* __restrict__ scb,
                                                    const double* __restrict__ Hb, double* __restrict__ dpart)
{
    // one workgroup per (shape b, node i), over chunks of 32 nodes j:
    // Gbar[j][t] = sum_s q_s(j) Hbar[s][t] on FP64 MFMA 16x16x4 (A = Q, j x s;
    // B = Hbar, s x t; wave w owns the j tile w / 2 and the t tile w % 2),
    // then each lane adds Gbar G_1 at its 4 output elements.  The Y and sin /
    // cos omega k_j of those elements and the basis entries of Qs in the next
    // chunk are loaded while the current chunk's products run.
    __shared__ double Hs[32][32];   // [s][t]
    __shared__ double Qs[32][32];   // [s][j in chunk]
    __shared__ double red[256];
    const int b = blockIdx.x % nb, i = blockIdx.x / nb;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int jm = 16 * (w >> 1), tn = 16 * (w & 1);
    const int t = tn + (lane & 15);   // the lane's output column; rows jm + 4 v + lane / 16
    const double ki = kn[i];
    double sa, ca;   // sin, cos of omega k_i + phi
    sincos(theta[2 * b] * ki + theta[2 * b + 1], &sa, &ca);
    const double* const sc = scb + 2 * (size_t)b * N;
    for (int e = tid; e < 32 * 32; e += 256) {
        const int ss = e >> 5, tt = e & 31;
        Hs[ss][tt] = (ss < R && tt < R) ? Hb[(((size_t)b * N) + i) * R * R + ss * R + tt] : 0.0;
    }
    T yv[4][2];
    double sv[4], cv[4], qn[4];
    // the lane's elements (j0 + jm + 4 v + lane / 16, t) of a chunk and the
    // thread's 4 entries of its Qs (nodes past N: the last node; columns past
    // R: the last column; all unused)
    auto fetch = [&](int j0) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int j = min(j0 + jm + 4 * v + (lane >> 4), N - 1), tt = min(t, R - 1);
            const T* y = Y + (size_t)rowof[pair_row(i, j, N)] * ldy + (size_t)b * 2 * R + tt;
            yv[v][0] = y[0];
            yv[v][1] = y[R];
            sv[v] = sc[2 * j];
            cv[v] = sc[2 * j + 1];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = tid + 256 * k, ss = e >> 5, jj = e & 31;
            qn[k] = q[(size_t)min(j0 + jj, N - 1) * ldq + min(ss, R - 1)];
        }
    };
    double pacc = 0.0;
    fetch(0);
    for (int j0 = 0; j0 < N; j0 += 32) {
        __syncthreads();   // Hs written / the previous chunk is done with Qs
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = tid + 256 * k, ss = e >> 5, jj = e & 31;
            Qs[ss][jj] = (ss < R && j0 + jj < N) ? qn[k] : 0.0;
        }
        __syncthreads();
        // this chunk's element data, then the next chunk's loads
        T yc[4][2];
        double scs[4], ccs[4];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            yc[v][0] = yv[v][0];
            yc[v][1] = yv[v][1];
            scs[v] = sv[v];
            ccs[v] = cv[v];
        }
        fetch(min(j0 + 32, N - 1));
        d4 gb = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
            const int ss = 4 * ks + (lane >> 4);
            gb = __builtin_amdgcn_mfma_f64_16x16x4f64(Qs[ss][jm + (lane & 15)], Hs[ss][t], gb, 0, 0, 0);
        }
        if (t < R)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                if (j0 + jm + 4 * v + (lane >> 4) < N) {
                    const double sA = sa * ccs[v] + ca * scs[v], cA = ca * ccs[v] - sa * scs[v];
                    const double yr = (double)yc[v][0], ys = (double)yc[v][1];
                    pacc += gb[v] * (cA * yr - sA * ys);
                }
            }
    }
    red[tid] = pacc;
    __syncthreads();
    for (int wd = 128; wd > 0; wd >>= 1) {
        if (tid < wd) red[tid] += red[tid + wd];
        __syncthreads();
    }
    if (tid == 0) {
        dpart[((size_t)b * N + i) * 2] = ki * red[0];
        dpart[((size_t)b * N + i) * 2 + 1] = red[0] * (1.0 / 3.0);
    }
}

// dtheta[b][x] = sum_i dpart[b][i][x]: one wave per (b, x), lane l summing
// i = l, l + 64, ..., then a butterfly over the lanes (a fixed order)
__global__ __launch_bounds__(64) void proj_sum_parts(const double* __restrict__ dpart, int N, int nb,
                                                     double* __restrict__ out)
{
    const int e = blockIdx.x, b = e / 2, x = e % 2, lane = threadIdx.x;
    double s = 0.0;
    for (int i = lane; i < N; i += 64) s += dpart[((size_t)b * N + i) * 2 + x];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
    if (lane == 0) out[e] = s;
}

// ---- host side ---------------------------------------------------------------

static int ensure_W(mcmb_ctx* c, int prec)
{
    PointSet& K = c->ks;
    const int N = K.lmax + 1, Kp = (N + PK - 1) / PK * PK;
    if (c->w_ready == prec + 1) return MCMB_OK;
    const long M = (long)N * (N + 1) / 2;
    // The pair rows i <= j in an order that groups rows of like nonzero k
    // range: W_ijk is nonzero only for k_k in [k_j - k_i, k_i + k_j] (the
    // triangle), so with the rows sorted by (upper end / PK, lower end) most
    // K chunks of a PT-row tile are zero throughout, and the GEMMs skip them
    // (43% of the chunks remain at 300 nodes, against 87% in i-major pair
    // order).  Each row's k-sum is unchanged, so Y holds the same values in
    // W's row order; wrow[pair_row(i, j)] is the row of (i, j).
    const std::vector<double>& kn = K.kn;
    std::vector<int2> rowij;
    std::vector<long> key;
    rowij.reserve(M);
    key.reserve(M);
    for (int i = 0; i < N; ++i)
        for (int j = i; j < N; ++j) {
            const long lo = std::lower_bound(kn.begin(), kn.end(), kn[j] - kn[i]) - kn.begin();
            const long hi = std::upper_bound(kn.begin(), kn.end(), kn[i] + kn[j]) - kn.begin();
            key.push_back((hi / PK) * (N + 1) + lo);
            rowij.push_back(make_int2(i, j));
        }
    std::vector<long> ord(M);
    for (long r = 0; r < M; ++r) ord[r] = r;
    std::stable_sort(ord.begin(), ord.end(), [&](long a, long b) { return key[a] < key[b]; });
    std::vector<int2> rows(M);
    std::vector<int> wrow(M);
    for (long r = 0; r < M; ++r) {
        rows[r] = rowij[ord[r]];
        wrow[ord[r]] = (int)r;   // rowij is in pair_row order
    }
    DevBuf d_ij;
    MCMB_TRY(d_ij.alloc(sizeof(int2) * (size_t)M));
    MCMB_HIP(hipMemcpyAsync(d_ij.p, rows.data(), sizeof(int2) * (size_t)M, hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(c->d_wrow.alloc(sizeof(int) * (size_t)M));
    MCMB_HIP(hipMemcpyAsync(c->d_wrow.p, wrow.data(), sizeof(int) * (size_t)M, hipMemcpyHostToDevice, c->stream));
    const unsigned ntile = (unsigned)((M + PT - 1) / PT);
    MCMB_TRY(c->d_wtile.alloc(sizeof(int2) * ntile));
    size_t tot = (size_t)M * Kp;
    MCMB_TRY(launch_guard("projection weight cube", (tot + 255) / 256, 256));
    if (prec == MCMB_FP64) {
        MCMB_TRY(c->d_W.alloc(tot * sizeof(double)));
        hipLaunchKernelGGL(proj_build_W<double>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream,
                           K.d_kn.as<double>(), K.d_kw.as<double>(), N, Kp, d_ij.as<int2>(), M,
                           c->d_W.as<double>());
        hipLaunchKernelGGL(proj_w_tiles<double>, dim3(ntile), dim3(256), 0, c->stream,
                           (const double*)c->d_W.as<double>(), M, Kp, c->d_wtile.as<int2>());
    } else {
        MCMB_TRY(c->d_W.alloc(tot * sizeof(float)));
        hipLaunchKernelGGL(proj_build_W<float>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream,
                           K.d_kn.as<double>(), K.d_kw.as<double>(), N, Kp, d_ij.as<int2>(), M,
                           c->d_W.as<float>());
        hipLaunchKernelGGL(proj_w_tiles<float>, dim3(ntile), dim3(256), 0, c->stream,
                           (const float*)c->d_W.as<float>(), M, Kp, c->d_wtile.as<int2>());
    }
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));   // the pair table goes out of scope
    c->w_ready = prec + 1;
    return MCMB_OK;
}

// Leading dimension of stage 1's Z and Y for a batch of nb shapes (4R
// columns per shape, rounded up to whole GEMM tiles).
static int stage1_ld(int nb, int R)
{
    return (nb * 2 * R + PT - 1) / PT * PT;
}

// FP32 stage-1 GEMM C[m][ldc] = A[m][Kp] B[Kp][ldb] (B zero-padded to ldb
// columns and Kp rows, A to Kp columns): 128 x 128 tiles, 4 waves of
// 64 x 64 on v_mfma_f32_16x16x4f32, K chunks of 16 in double-buffered LDS,
// over the tile's chunks ktiles[tile] = (first, count) only (A is zero in
// the others).  Operands come in as 16-byte loads two chunks ahead of their
// use (two register sets), A stored [m][k] and B [k][n] with row strides (20
// and 144 floats) that keep both the 16-byte stores and the MFMA operand
// reads free of bank conflicts.  The chunks run in K order.
constexpr int PF_LDA = PK + 4, PF_LDB = PT + 16;
__global__ __launch_bounds__(256, 2) void proj_gemm_f32(const float* __restrict__ A, const float* __restrict__ B,
                                                        int M, int Kp, int ldb, float* __restrict__ C, int ldc,
                                                        const int2* __restrict__ ktiles)
{
    __shared__ __align__(16) float As[2][PT][PF_LDA];
    __shared__ __align__(16) float Bs[2][PK][PF_LDB];
    const int m0 = blockIdx.y * PT, n0 = blockIdx.x * PT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    f4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (f4){0.f, 0.f, 0.f, 0.f};
    // a chunk is 512 float4 of A (row r = e / 4, k = 4 (e % 4)) and 512 of B
    // (k = e / 32, n = 4 (e % 32)), e = tid + 256 u
    const f4 z4 = {0.f, 0.f, 0.f, 0.f};
    auto gload = [&](int k0, f4 (&ra)[2], f4 (&rb)[2]) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + 256 * u;
            const int r = e >> 2, kq = e & 3, gm = m0 + r;
            ra[u] = gm < M ? *reinterpret_cast<const f4*>(A + (size_t)gm * Kp + k0 + 4 * kq) : z4;
            const int kb = e >> 5, c4 = e & 31;
            rb[u] = *reinterpret_cast<const f4*>(B + (size_t)(k0 + kb) * ldb + n0 + 4 * c4);
        }
    };
    auto sstore = [&](int buf, const f4 (&ra)[2], const f4 (&rb)[2]) {
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int e = tid + 256 * u;
            *reinterpret_cast<f4*>(&As[buf][e >> 2][4 * (e & 3)]) = ra[u];
            *reinterpret_cast<f4*>(&Bs[buf][e >> 5][4 * (e & 31)]) = rb[u];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < PK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            float av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][wm * 64 + a * 16 + (lane & 15)][kk];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[buf][kk][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
    };
    // the tile's chunks c0 .. c0 + nk - 1 (the rest of W's rows here are
    // zero); chunk kc is in LDS buffer kc & 1, chunk kc + 1 in registers and
    // chunk kc + 2 loading; loads past the last chunk repeat it (unused), so
    // every load is unconditional
    const int2 kr = ktiles[blockIdx.y];
    const int c0 = kr.x, nk = kr.y;
    f4 ra0[2], rb0[2], ra1[2], rb1[2];
    gload(c0 * PK, ra0, rb0);
    gload((c0 + min(1, nk - 1)) * PK, ra1, rb1);
    sstore(0, ra0, rb0);
    __syncthreads();
    for (int kc = 0;; kc += 2) {
        gload((c0 + min(kc + 2, nk - 1)) * PK, ra0, rb0);
        compute(0);
        sstore(1, ra1, rb1);
        __syncthreads();
        if (kc + 1 >= nk) break;
        gload((c0 + min(kc + 3, nk - 1)) * PK, ra1, rb1);
        compute(1);
        sstore(0, ra0, rb0);
        __syncthreads();
        if (kc + 2 >= nk) break;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm * 64 + a * 16 + MF<float>::row(v, lane);
                const int gn = n0 + wn * 64 + b * 16 + (lane & 15);
                if (gm < M) C[(size_t)gm * ldc + gn] = acc[a][b][v];
            }
}

// FP64 stage-1 GEMM, the same scheme with 16-byte (two-double) operand
// loads two chunks ahead: a chunk is 1024 double2 of A (row r = e / 8, k = 2
// (e % 8)) and 1024 of B (k = e / 64, n = 2 (e % 64)), e = tid + 256 u, u <
// 4.  A is stored [m][k] with a row stride of 18 doubles (16-byte aligned
// rows), B [k][n] with 136.  It issues the MFMA sequence of proj_gemm<double>
// (same operands, same K order).
constexpr int PD_LDA = PK + 2, PD_LDB = PT + 8;
typedef double d2v __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256, 2) void proj_gemm_f64(const double* __restrict__ A, const double* __restrict__ B,
                                                        int M, int Kp, int ldb, double* __restrict__ C, int ldc,
                                                        const int2* __restrict__ ktiles)
{
    __shared__ __align__(16) double As[2][PT][PD_LDA];
    __shared__ __align__(16) double Bs[2][PK][PD_LDB];
    const int m0 = blockIdx.y * PT, n0 = blockIdx.x * PT;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = (d4){0.0, 0.0, 0.0, 0.0};
    const d2v z2 = {0.0, 0.0};
    auto gload = [&](int k0, d2v (&ra)[4], d2v (&rb)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int r = e >> 3, kq = e & 7, gm = m0 + r;
            ra[u] = gm < M ? *reinterpret_cast<const d2v*>(A + (size_t)gm * Kp + k0 + 2 * kq) : z2;
            const int kb = e >> 6, c2 = e & 63;
            rb[u] = *reinterpret_cast<const d2v*>(B + (size_t)(k0 + kb) * ldb + n0 + 2 * c2);
        }
    };
    auto sstore = [&](int buf, const d2v (&ra)[4], const d2v (&rb)[4]) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            *reinterpret_cast<d2v*>(&As[buf][e >> 3][2 * (e & 7)]) = ra[u];
            *reinterpret_cast<d2v*>(&Bs[buf][e >> 6][2 * (e & 63)]) = rb[u];
        }
    };
    auto compute = [&](int buf) {
#pragma unroll
        for (int ks = 0; ks < PK / 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][wm * 64 + a * 16 + (lane & 15)][kk];
#pragma unroll
            for (int b = 0; b < 4; ++b) bv[b] = Bs[buf][kk][wn * 64 + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0, 0);
        }
    };
    // the tile's chunks c0 .. c0 + nk - 1, as proj_gemm_f32
    const int2 kr = ktiles[blockIdx.y];
    const int c0 = kr.x, nk = kr.y;
    d2v ra0[4], rb0[4], ra1[4], rb1[4];
    gload(c0 * PK, ra0, rb0);
    gload((c0 + min(1, nk - 1)) * PK, ra1, rb1);
    sstore(0, ra0, rb0);
    __syncthreads();
    for (int kc = 0;; kc += 2) {
        gload((c0 + min(kc + 2, nk - 1)) * PK, ra0, rb0);
        compute(0);
        sstore(1, ra1, rb1);
        __syncthreads();
        if (kc + 1 >= nk) break;
        gload((c0 + min(kc + 3, nk - 1)) * PK, ra1, rb1);
        compute(1);
        sstore(0, ra0, rb0);
        __syncthreads();
        if (kc + 2 >= nk) break;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm * 64 + a * 16 + MF<double>::row(v, lane);
                const int gn = n0 + wn * 64 + b * 16 + (lane & 15);
                if (gm < M) C[(size_t)gm * ldc + gn] = acc[a][b][v];
            }
}

template <class T>
static int run_stage1(mcmb_ctx* c, const double* d_theta, int nb, int& ldy, DevBuf& Yb)
{
    DevBuf& Zb = c->pz;
    PointSet& K = c->ks;
    const int N = K.lmax + 1, R = K.R, Kp = (N + PK - 1) / PK * PK;
    const int ldz = stage1_ld(nb, R);
    ldy = ldz;
    MCMB_TRY(Zb.alloc(sizeof(T) * (size_t)Kp * ldz));
    const long M = (long)N * (N + 1) / 2;   // pair rows i <= j
    MCMB_TRY(Yb.alloc(sizeof(T) * (size_t)M * ldy));
    size_t ze = (size_t)Kp * ldz;
    hipLaunchKernelGGL(proj_build_Z<T>, dim3((unsigned)((ze + 255) / 256)), dim3(256), 0, c->stream,
                       K.d_kn.as<double>(), K.d_q.as<double>(), K.ldq, N, R, d_theta, nb, Kp, ldz, Zb.as<T>());
    dim3 grid((unsigned)(ldz / PT), (unsigned)((M + PT - 1) / PT));
    if (std::is_same<T, float>::value)
        hipLaunchKernelGGL(proj_gemm_f32, grid, dim3(256), 0, c->stream, (const float*)c->d_W.as<T>(),
                           (const float*)Zb.as<T>(), (int)M, Kp, ldz, (float*)Yb.as<T>(), ldy,
                           (const int2*)c->d_wtile.as<int2>());
    else
        hipLaunchKernelGGL(proj_gemm_f64, grid, dim3(256), 0, c->stream, (const double*)c->d_W.as<T>(),
                           (const double*)Zb.as<T>(), (int)M, Kp, ldz, (double*)Yb.as<T>(), ldy,
                           (const int2*)c->d_wtile.as<int2>());
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// sin / cos of omega_b k_j for the nb shapes of a batch into c->psc
static int sincos_table(mcmb_ctx* c, const double* d_theta, int nb)
{
    const int N = c->ks.lmax + 1;
    MCMB_TRY(c->psc.alloc(sizeof(double) * 2 * (size_t)PBATCH * N));
    hipLaunchKernelGGL(proj_sincos_tab, dim3((unsigned)((nb * N + 255) / 256)), dim3(256), 0, c->stream,
                       c->ks.d_kn.as<double>(), d_theta, N, nb, c->psc.as<double>());
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// Y is stage 1's product buffer (the context's batch buffer, or a slot of
// the forward cache).
template <class T>
static int project_fwd_batch(mcmb_ctx* c, const double* d_theta, int nb, double* d_alpha /* [nb][n] */,
                             double* d_dalpha /* [nb][2][n] */, DevBuf& Yb)
{
    PointSet& K = c->ks;
    const int N = K.lmax + 1, R = K.R, n = K.n;
    DevBuf& Hb = c->ph;
    int ldy = 0;
    MCMB_TRY(run_stage1<T>(c, d_theta, nb, ldy, Yb));
    MCMB_TRY(Hb.alloc(sizeof(double) * (size_t)PBATCH * 2 * N * R * R));
    MCMB_TRY(sincos_table(c, d_theta, nb));
    hipLaunchKernelGGL(proj_stage2<T>, dim3((unsigned)(nb * N)), dim3(256), 0, c->stream, Yb.as<T>(), ldy,
                       (const int*)c->d_wrow.as<int>(), K.d_kn.as<double>(), K.d_q.as<double>(), K.ldq, N, R, d_theta, nb,
                       (const double*)c->psc.as<double>(), Hb.as<double>());
    MCMB_TRY(c->pdm.alloc(sizeof(double) * (size_t)PBATCH * 3 * R * R * R));
    hipLaunchKernelGGL(proj_stage3, dim3((unsigned)(nb * 2 * ((R * R + 15) / 16))), dim3(64), 0, c->stream,
                       (const double*)Hb.as<double>(), K.d_q.as<double>(), K.ldq, K.d_kn.as<double>(), N, R,
                       c->pdm.as<double>());
    const int ne = nb * 3 * n;
    hipLaunchKernelGGL(proj_stage3_modes, dim3((ne + 255) / 256), dim3(256), 0, c->stream,
                       (const double*)c->pdm.as<double>(), R, K.d_modes.as<int>(), n, nb, d_alpha, d_dalpha);
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// Ycached: stage 1's product of this batch from the forward cache, or null to
// compute it.
template <class T>
static int project_bwd_batch(mcmb_ctx* c, const double* d_theta, int nb, const double* d_gbar, const int* d_slot,
                             double* d_dtheta, DevBuf* Ycached)
{
    PointSet& K = c->ks;
    const int N = K.lmax + 1, R = K.R, n = K.n;
    DevBuf& Hb = c->ph;
    DevBuf& Dp = c->pd;
    // a cached product is this batch's own: same shape count, so the same
    // leading dimension; anything smaller than that is recomputed
    const long M = (long)N * (N + 1) / 2;
    int ldy = stage1_ld(nb, R);
    if (Ycached && Ycached->bytes < sizeof(T) * (size_t)M * ldy) Ycached = nullptr;
    DevBuf& Yb = Ycached ? *Ycached : c->py;
    if (!Ycached) MCMB_TRY(run_stage1<T>(c, d_theta, nb, ldy, Yb));
    MCMB_TRY(Hb.alloc(sizeof(double) * (size_t)PBATCH * N * R * R));
    MCMB_TRY(Dp.alloc(sizeof(double) * (size_t)PBATCH * N * 2));
    const unsigned hblocks = (unsigned)(nb * ((R * R + 15) / 16));
    hipLaunchKernelGGL(proj_hbar, dim3(hblocks), dim3(64), 0, c->stream, d_gbar, d_slot, K.d_q.as<double>(), K.ldq,
                       N, R, n, nb, Hb.as<double>());
    MCMB_TRY(sincos_table(c, d_theta, nb));
    hipLaunchKernelGGL(proj_adjoint<T>, dim3((unsigned)(nb * N)), dim3(256), 0, c->stream, Yb.as<T>(), ldy,
                       (const int*)c->d_wrow.as<int>(), K.d_kn.as<double>(), K.d_q.as<double>(), K.ldq, N, R, d_theta, nb,
                       (const double*)c->psc.as<double>(), Hb.as<double>(), Dp.as<double>());
    hipLaunchKernelGGL(proj_sum_parts, dim3((unsigned)(nb * 2)), dim3(64), 0, c->stream, Dp.as<double>(), N, nb,
                       d_dtheta);
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// The forward cache holds every batch's stage-1 product when they fit in
// PCACHE_BYTES together.
constexpr size_t PCACHE_BYTES = 24ull << 30;

static size_t ybytes(const mcmb_ctx* c, int prec)
{
    const PointSet& K = c->ks;
    const int N = K.lmax + 1, R = K.R;
    const long M = (long)N * (N + 1) / 2;
    return (prec == MCMB_FP64 ? sizeof(double) : sizeof(float)) * (size_t)M * stage1_ld(PBATCH, R);
}

}  // namespace mcmb

using namespace mcmb;

extern "C" {

int mcmb_kspace_set(mcmb_ctx* c, int N, const double* kn, const double* kw, int ldq, const double* q, int nmodes,
                    const int* order)
{
    if (!c || !kn || !kw || !q || !order) return fail(MCMB_ERR_ARG, "mcmb_kspace_set: null argument");
    if (N < 2 || N > 4096) return fail(MCMB_ERR_ARG, "mcmb_kspace_set: N %d not in [2,4096]", N);
    if (ldq < 1 || ldq > 32) return fail(MCMB_ERR_ARG, "mcmb_kspace_set: ldq %d not in [1,32]", ldq);
    if (nmodes < 1 || nmodes > 65536) return fail(MCMB_ERR_ARG, "mcmb_kspace_set: nmodes %d", nmodes);
    for (int a = 1; a < N; ++a)
        if (!(kn[a] > kn[a - 1]) || !(kn[0] > 0.0))
            return fail(MCMB_ERR_ARG, "mcmb_kspace_set: nodes must be positive and strictly ascending");
    int R = 0;
    for (int m = 0; m < nmodes; ++m)
        for (int a = 0; a < 3; ++a) {
            int p = order[3 * m + a];
            if (p < 0 || p >= ldq)
                return fail(MCMB_ERR_ORDER, "mcmb_kspace_set: order[%d][%d] = %d outside [0,%d)", m, a, p, ldq);
            if (p + 1 > R) R = p + 1;
        }
    MCMB_TRY(set_device(c));
    PointSet& K = c->ks;
    K.ready = false;
    K.kspace = true;
    K.kstep = 1;
    K.lmin = 0;
    K.lmax = N - 1;
    K.ldq = ldq;
    K.n = nmodes;
    K.R = R;
    K.order.assign(order, order + 3 * (size_t)nmodes);
    K.q.assign(q, q + (size_t)N * ldq);
    K.kn.assign(kn, kn + N);
    K.kw.assign(kw, kw + N);
    // ordered masked points a <= b <= c, k_c <= k_a + k_b (nodes ascending)
    K.rows.clear();
    K.npoints = 0;
    for (int a = 0; a < N; ++a)
        for (int b = a; b < N; ++b) {
            int cmax = b;
            while (cmax + 1 < N && kn[cmax + 1] <= kn[a] + kn[b]) ++cmax;
            K.rows.push_back(RowDesc{a, b, b, cmax - b + 1});
            K.npoints += cmax - b + 1;
        }
    MCMB_TRY(K.d_kn.alloc(sizeof(double) * N));
    MCMB_TRY(K.d_kw.alloc(sizeof(double) * N));
    MCMB_HIP(hipMemcpyAsync(K.d_kn.p, kn, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(K.d_kw.p, kw, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(upload_pointset(c, K, K.order));
    c->w_ready = 0;
    K.ready = true;
    K.generation = next_generation();
    return MCMB_OK;
}

long mcmb_kspace_npoints(mcmb_ctx* c)
{
    if (!c || !c->ks.ready) return fail(MCMB_ERR_STATE, "mcmb_kspace_npoints: no k-space setup");
    return c->ks.npoints;
}

int mcmb_kgram(mcmb_ctx* c, int precision, double* gamma_host, double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_kgram: null context");
    if (!c->ks.ready) return fail(MCMB_ERR_STATE, "mcmb_kgram: call mcmb_kspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_MIXED) return fail(MCMB_ERR_ARG, "mcmb_kgram: precision %d", precision);
    MCMB_TRY(set_device(c));
    const int n = c->ks.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_TRY(gram_tet_launch(c, c->ks, precision, dst, 0, (int)c->ks.rows.size()));
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_project_fwd(mcmb_ctx* c, int precision, int ntheta, const double* theta, double* alpha, double* dalpha)
{
    if (!c || !theta || !alpha || ntheta < 1) return fail(MCMB_ERR_ARG, "mcmb_project_fwd: bad argument");
    if (!c->ks.ready) return fail(MCMB_ERR_STATE, "mcmb_project_fwd: call mcmb_kspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_FP32) return fail(MCMB_ERR_ARG, "mcmb_project_fwd: precision %d", precision);
    MCMB_TRY(set_device(c));
    const int n = c->ks.n;
    MCMB_TRY(ensure_W(c, precision));
    const int nbat = (ntheta + PBATCH - 1) / PBATCH;
    // keep each batch's stage-1 product for a backward call on the same shapes
    const bool cache = (size_t)nbat * ybytes(c, precision) <= PCACHE_BYTES;
    c->ycache_gen = -1;
    if (cache && (int)c->ycache.size() < nbat) c->ycache.resize(nbat);
    // device arrays of the call: theta [ntheta][2], alpha [ntheta][n], dalpha [ntheta][2][n]
    MCMB_TRY(c->pio.alloc(sizeof(double) * (2 + 3 * (size_t)n) * ntheta));
    double* const dth = c->pio.as<double>();
    double* const dal = dth + 2 * (size_t)ntheta;
    double* const ddal = dal + (size_t)ntheta * n;
    // Each batch's results go to the host while later batches compute: once
    // the batch's kernels are done, the copy stream moves its alpha (and
    // dalpha) into slot bi % nslot of the pinned staging, and host threads
    // move the slots on to the caller's arrays in batch order, so only the
    // last batch's transfer follows the GPU work.
    const int nout = dalpha ? 3 : 1;
    const size_t slot_bytes = sizeof(double) * (size_t)PBATCH * n * nout;
    const int nslot = (int)std::min<size_t>(PSLOTS, 2 * STAGE_BYTES / slot_bytes);
    const char* cenv = getenv("MCMB_HOST_COPY");
    const bool staged = nslot >= 1 && !(cenv && !strcmp(cenv, "direct"));
    if (staged) MCMB_TRY(copy_stream_ready(c));
    const int nth = std::min(8, host_threads());
    auto drain = [&](int bi) -> int {
        MCMB_HIP(hipEventSynchronize(c->slot_ev[1][bi % nslot]));
        const int b0 = bi * PBATCH, nb = std::min(PBATCH, ntheta - b0);
        const double* sl = reinterpret_cast<const double*>(static_cast<char*>(c->stage) + (bi % nslot) * slot_bytes);
        const size_t na = (size_t)nb * n;
        run_threads(nth, [&](int t) {
            const size_t a = na * t / nth, b = na * (t + 1) / nth;
            memcpy(alpha + (size_t)b0 * n + a, sl + a, sizeof(double) * (b - a));
            if (dalpha) memcpy(dalpha + (size_t)b0 * 2 * n + 2 * a, sl + na + 2 * a, sizeof(double) * 2 * (b - a));
        });
        return MCMB_OK;
    };
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    MCMB_HIP(hipMemcpyAsync(dth, theta, sizeof(double) * 2 * (size_t)ntheta, hipMemcpyHostToDevice, c->stream));
    for (int b0 = 0, bi = 0; b0 < ntheta; b0 += PBATCH, ++bi) {
        const int nb = std::min(PBATCH, ntheta - b0);
        DevBuf& Yb = cache ? c->ycache[bi] : c->py;
        double* d_al = dal + (size_t)b0 * n;
        double* d_dal = ddal + (size_t)b0 * 2 * n;
        if (precision == MCMB_FP64)
            MCMB_TRY(project_fwd_batch<double>(c, dth + 2 * b0, nb, d_al, d_dal, Yb));
        else
            MCMB_TRY(project_fwd_batch<float>(c, dth + 2 * b0, nb, d_al, d_dal, Yb));
        if (staged) {
            if (bi >= nslot) MCMB_TRY(drain(bi - nslot));   // the slot's previous batch
            const int k = bi % nslot;
            char* sl = static_cast<char*>(c->stage) + k * slot_bytes;
            MCMB_HIP(hipEventRecord(c->slot_ev[0][k], c->stream));
            MCMB_HIP(hipStreamWaitEvent(c->cstream, c->slot_ev[0][k], 0));
            MCMB_HIP(hipMemcpyAsync(sl, d_al, sizeof(double) * (size_t)nb * n, hipMemcpyDeviceToHost, c->cstream));
            if (dalpha)
                MCMB_HIP(hipMemcpyAsync(sl + sizeof(double) * (size_t)nb * n, d_dal, sizeof(double) * 2 * (size_t)nb * n,
                                        hipMemcpyDeviceToHost, c->cstream));
            MCMB_HIP(hipEventRecord(c->slot_ev[1][k], c->cstream));
        }
    }
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    if (staged) {
        for (int bi = std::max(0, nbat - nslot); bi < nbat; ++bi) MCMB_TRY(drain(bi));
    } else {
        // the device arrays already have the caller's layout: one contiguous copy each
        MCMB_HIP(hipMemcpyAsync(alpha, dal, sizeof(double) * (size_t)ntheta * n, hipMemcpyDeviceToHost, c->stream));
        if (dalpha)
            MCMB_HIP(hipMemcpyAsync(dalpha, ddal, sizeof(double) * (size_t)ntheta * 2 * n, hipMemcpyDeviceToHost,
                                    c->stream));
    }
    MCMB_HIP(hipStreamSynchronize(c->stream));
    if (cache) {
        c->ycache_theta.assign(theta, theta + 2 * (size_t)ntheta);
        c->ycache_prec = precision;
        c->ycache_gen = c->ks.generation;
    }
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = nbat;
    return MCMB_OK;
}

int mcmb_project_bwd(mcmb_ctx* c, int precision, int ntheta, const double* theta, const double* gbar, double* dtheta)
{
    if (!c || !theta || !gbar || !dtheta || ntheta < 1) return fail(MCMB_ERR_ARG, "mcmb_project_bwd: bad argument");
    if (!c->ks.ready) return fail(MCMB_ERR_STATE, "mcmb_project_bwd: call mcmb_kspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_FP32) return fail(MCMB_ERR_ARG, "mcmb_project_bwd: precision %d", precision);
    MCMB_TRY(set_device(c));
    PointSet& K = c->ks;
    const int n = K.n, R = K.R;
    MCMB_TRY(ensure_W(c, precision));
    // slot[r][s][t] = mode index of the canonical triple (r,s,t) = order[m]
    std::vector<int> slot((size_t)R * R * R, -1);
    for (int m = 0; m < n; ++m) {
        int r = K.order[3 * m], s = K.order[3 * m + 1], t = K.order[3 * m + 2];
        int& sl = slot[((size_t)r * R + s) * R + t];
        if (sl >= 0) return fail(MCMB_ERR_ORDER, "mcmb_project_bwd: repeated mode (%d,%d,%d)", r, s, t);
        sl = m;
    }
    // device arrays of the call: theta [ntheta][2], gbar [ntheta][n], dtheta
    // [ntheta][2], then the slot table
    MCMB_TRY(c->pio.alloc(sizeof(double) * (4 + (size_t)n) * ntheta + sizeof(int) * slot.size()));
    double* const dth = c->pio.as<double>();
    double* const dg = dth + 2 * (size_t)ntheta;
    double* const dd = dg + (size_t)ntheta * n;
    int* const dslot = reinterpret_cast<int*>(dd + 2 * (size_t)ntheta);
    // the last forward's stage-1 products, when it ran on exactly these shapes
    const bool cached = c->ycache_gen == K.generation && c->ycache_prec == precision &&
                        c->ycache_theta.size() == 2 * (size_t)ntheta &&
                        !memcmp(c->ycache_theta.data(), theta, sizeof(double) * 2 * (size_t)ntheta);
    // gbar goes up batch by batch through slots of the pinned staging: host
    // threads fill slot bi % nslot while the GPU works on earlier batches
    const size_t slot_bytes = sizeof(double) * (size_t)PBATCH * n;
    const int nslot = (int)std::min<size_t>(PSLOTS, 2 * STAGE_BYTES / slot_bytes);
    const char* cenv = getenv("MCMB_HOST_COPY");
    const bool staged = nslot >= 1 && !(cenv && !strcmp(cenv, "direct"));
    if (staged) MCMB_TRY(copy_stream_ready(c));
    const int nth = std::min(8, host_threads());
    MCMB_HIP(hipMemcpyAsync(dslot, slot.data(), sizeof(int) * slot.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    MCMB_HIP(hipMemcpyAsync(dth, theta, sizeof(double) * 2 * (size_t)ntheta, hipMemcpyHostToDevice, c->stream));
    if (!staged)
        MCMB_HIP(hipMemcpyAsync(dg, gbar, sizeof(double) * (size_t)ntheta * n, hipMemcpyHostToDevice, c->stream));
    for (int b0 = 0, bi = 0; b0 < ntheta; b0 += PBATCH, ++bi) {
        const int nb = std::min(PBATCH, ntheta - b0);
        DevBuf* Yc = cached ? &c->ycache[bi] : nullptr;
        const double* th = dth + 2 * b0;
        double* gb = dg + (size_t)b0 * n;
        double* dt = dd + 2 * b0;
        if (staged) {
            const int k = bi % nslot;
            // the slot's previous upload has been read
            if (bi >= nslot) MCMB_HIP(hipEventSynchronize(c->slot_ev[0][k]));
            double* sl = reinterpret_cast<double*>(static_cast<char*>(c->stage) + k * slot_bytes);
            const size_t na = (size_t)nb * n;
            run_threads(nth, [&](int t) {
                const size_t a = na * t / nth, b = na * (t + 1) / nth;
                memcpy(sl + a, gbar + (size_t)b0 * n + a, sizeof(double) * (b - a));
            });
            MCMB_HIP(hipMemcpyAsync(gb, sl, sizeof(double) * na, hipMemcpyHostToDevice, c->stream));
            MCMB_HIP(hipEventRecord(c->slot_ev[0][k], c->stream));
        }
        if (precision == MCMB_FP64)
            MCMB_TRY(project_bwd_batch<double>(c, th, nb, gb, dslot, dt, Yc));
        else
            MCMB_TRY(project_bwd_batch<float>(c, th, nb, gb, dslot, dt, Yc));
    }
    MCMB_HIP(hipMemcpyAsync(dtheta, dd, sizeof(double) * 2 * (size_t)ntheta, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = (ntheta + PBATCH - 1) / PBATCH;
    return MCMB_OK;
}

}  // extern "C"

// ---- Gamma post-processing (gamma_mpi.c:260-295, 801-840) -------------------
// orthoinv = lambda^T lambda, Gamma' = orthoinv Gamma, FP64 MFMA products
// (proj_gemm<double>: A n x Kp, B Kp x n, both zero-padded in K).
__global__ void la_pad(const double* __restrict__ src, int n, int trans, int rows, int cols,
                       double* __restrict__ dst)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)rows * cols) return;
    const int r = (int)(e / cols), c = (int)(e % cols);
    dst[e] = (r < n && c < n) ? (trans ? src[(long)c * n + r] : src[(long)r * n + c]) : 0.0;
}

extern "C" int mcmb_gamma_post(mcmb_ctx* c, int n, const double* lambda, const double* gamma, double* orthoinv,
                               double* gamma_out)
{
    if (!c || n < 1 || !lambda || !gamma || !gamma_out) return fail(MCMB_ERR_ARG, "mcmb_gamma_post: bad arguments");
    MCMB_TRY(set_device(c));
    const int Kp = (n + PK - 1) / PK * PK;
    const size_t nn = (size_t)n * n, np = (size_t)n * Kp;
    DevBuf dL, dG, dA, dB, dC;
    MCMB_TRY(dL.alloc(sizeof(double) * nn));
    MCMB_TRY(dG.alloc(sizeof(double) * nn));
    MCMB_TRY(dA.alloc(sizeof(double) * np));
    MCMB_TRY(dB.alloc(sizeof(double) * np));
    MCMB_TRY(dC.alloc(sizeof(double) * nn));
    MCMB_TRY(copy_to_device(c, dL.p, lambda, sizeof(double) * nn));
    MCMB_TRY(copy_to_device(c, dG.p, gamma, sizeof(double) * nn));
    const unsigned gp = (unsigned)((np + 255) / 256);
    const dim3 grid((unsigned)((n + PT - 1) / PT), (unsigned)((n + PT - 1) / PT));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    // orthoinv[i][j] = sum_k lambda[k][i] lambda[k][j]
    hipLaunchKernelGGL(la_pad, dim3(gp), dim3(256), 0, c->stream, dL.as<double>(), n, 1, n, Kp, dA.as<double>());
    hipLaunchKernelGGL(la_pad, dim3(gp), dim3(256), 0, c->stream, dL.as<double>(), n, 0, Kp, n, dB.as<double>());
    hipLaunchKernelGGL(proj_gemm<double>, grid, dim3(256), 0, c->stream, dA.as<double>(), dB.as<double>(), n, Kp, n,
                       n, dC.as<double>(), n);
    // Gamma'[i][j] = sum_m orthoinv[i][m] Gamma[m][j]
    hipLaunchKernelGGL(la_pad, dim3(gp), dim3(256), 0, c->stream, dC.as<double>(), n, 0, n, Kp, dA.as<double>());
    hipLaunchKernelGGL(la_pad, dim3(gp), dim3(256), 0, c->stream, dG.as<double>(), n, 0, Kp, n, dB.as<double>());
    hipLaunchKernelGGL(proj_gemm<double>, grid, dim3(256), 0, c->stream, dA.as<double>(), dB.as<double>(), n, Kp, n,
                       n, dG.as<double>(), n);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    if (orthoinv) MCMB_TRY(copy_to_host(c, orthoinv, dC.p, sizeof(double) * nn));
    MCMB_TRY(copy_to_host(c, gamma_out, dG.p, sizeof(double) * nn));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 2;
    return MCMB_OK;
}
