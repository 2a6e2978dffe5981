// FP64 symmetric eigen-decomposition on the GPU (parallel cyclic Jacobi) and
// the late-side orthonormalisation with its PCA fallback
// (ortho_late_mpi.c:752-842; gsl_eigen_symmv + sort descending at
// ortho_prim_mpi.c:223-224 and ortho_late_mpi.c:775-776).
//
// Jacobi schedule.  n_e = n rounded up to even (an odd n gets one zero row and
// column at index n, which no rotation ever touches: its a_pq stay 0).  Each
// step rotates n_e/2 disjoint pairs at once.  The matrix is kept in a layout
// where the pairs of the current step are adjacent positions (2I, 2I+1); after
// the step every position x moves to rr_next(x), the circle method of a
// round-robin tournament (position 0 fixed, the others rotate), so that in
// n_e - 1 steps every pair of indices meets exactly once and the layout is
// back in the original order at the end of each sweep.  One kernel computes
// the n_e/2 rotations of a step, one kernel applies them: each thread rewrites
// one 2 x 2 block of A from its four inputs, A' = R_I^T A R_J, and one pair of
// eigenvector rows, writing both at their next-step positions (double
// buffered).  Rotations follow Golub & Van Loan (Matrix Computations, Alg.
// 8.5.1): tau = (a_qq - a_pp)/(2 a_pq), t = sign(tau)/(|tau| + sqrt(1+tau^2)),
// c = 1/sqrt(1+t^2), s = t c; a pair is skipped when |a_pq| <= 1e-16
// sqrt(|a_pp a_qq|) (relative accuracy of well-separated small eigenvalues) or
// |a_pq| <= 1e-17 ||A||_F (entries below the normwise backward error, which
// otherwise keep rotating among eigenvalues at the FP64 noise floor of a
// singular Gram).  Sweeps stop when a sweep applies no rotation or the
// off-diagonal norm is below max(1e-15, 2e-17 n) ||A||_F, inside the normwise
// bound n eps ||A|| of a Householder/QR eigensolver (eps = 1.1e-16); tighter
// stops only rotate inside clusters of eigenvalues at the FP64 noise floor,
// where Jacobi converges linearly.
#include "mcmb_internal.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace mcmb {

typedef double fd4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int rr_next(int x, int m)
{
    if (x == 0 || m == 1) return x;
    if (!(x & 1)) {
        const int i = x >> 1;
        return (i + 1 <= m - 1) ? 2 * (i + 1) : 2 * (m - 1) + 1;
    }
    const int i = (x - 1) >> 1;
    return i >= 1 ? 2 * (i - 1) + 1 : 2;
}

// C = D^-1 G D^-1 (norm[i] = sqrt G_ii) into the n_e x n_e buffer A (padding
// row/column zero), U = identity rows (U[x][i] = component i of the vector
// at layout position x; n_e rows of n).
__global__ void jac_init(const double* __restrict__ G, int n, int ne, int normalise, double* __restrict__ norm,
                         double* __restrict__ A, double* __restrict__ U)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)ne * ne;
    if (e < tot) {
        const int i = (int)(e / ne), j = (int)(e % ne);
        double v = 0.0;
        if (i < n && j < n) {
            v = G[(long)i * n + j];
            if (normalise) v /= sqrt(G[(long)i * n + i]) * sqrt(G[(long)j * n + j]);
        }
        A[e] = v;
    }
    if (e < (long)ne * n) {
        const int x = (int)(e / n), i = (int)(e % n);
        U[e] = (x == i) ? 1.0 : 0.0;
    }
    if (e < n && norm) norm[e] = sqrt(G[e * n + e]);
}

// Rotation of pair I = (2I, 2I+1) of the current layout: cs[I] = (c, s, t).
__global__ void jac_rot(const double* __restrict__ A, int ne, double abs_tol, double3* __restrict__ cs,
                        unsigned long long* __restrict__ nrot)
{
    const int I = blockIdx.x * blockDim.x + threadIdx.x;
    const int m = ne / 2;
    if (I >= m) return;
    const int p = 2 * I, q = p + 1;
    const double app = A[(long)p * ne + p], aqq = A[(long)q * ne + q], apq = A[(long)p * ne + q];
    double c = 1.0, s = 0.0, t = 0.0;
    if (fabs(apq) > abs_tol && fabs(apq) > 1e-16 * sqrt(fabs(app * aqq))) {
        const double tau = (aqq - app) / (2.0 * apq);
        t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        c = 1.0 / sqrt(1.0 + t * t);
        s = t * c;
        atomicAdd(nrot, 1ull);
    }
    cs[I] = make_double3(c, s, t);
}

// Apply the step's rotations: threads [0, m*m) rewrite 2 x 2 blocks of A,
// threads [m*m, m*m + m*n) rewrite pairs of U rows; outputs at the next
// step's positions.
__global__ __launch_bounds__(256) void jac_apply(const double* __restrict__ A, double* __restrict__ An,
                                                 const double* __restrict__ U, double* __restrict__ Un,
                                                 const double3* __restrict__ cs, int ne, int n)
{
    const int m = ne / 2;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long na = (long)m * m;
    if (e < na) {
        const int I = (int)(e / m), J = (int)(e % m);
        const double3 ri = cs[I], rj = cs[J];
        const long r0 = (long)(2 * I) * ne, r1 = r0 + ne;
        const double2 top = *reinterpret_cast<const double2*>(A + r0 + 2 * J);
        const double2 bot = *reinterpret_cast<const double2*>(A + r1 + 2 * J);
        double b00, b01, b10, b11;
        if (I == J) {
            const double apq = top.y;
            b00 = top.x - ri.z * apq;
            b11 = bot.y + ri.z * apq;
            b01 = b10 = 0.0;
            if (ri.y == 0.0) { b00 = top.x; b11 = bot.y; b01 = top.y; b10 = bot.x; }
        } else {
            const double x00 = top.x * rj.x - top.y * rj.y, x01 = top.x * rj.y + top.y * rj.x;
            const double x10 = bot.x * rj.x - bot.y * rj.y, x11 = bot.x * rj.y + bot.y * rj.x;
            b00 = ri.x * x00 - ri.y * x10;
            b01 = ri.x * x01 - ri.y * x11;
            b10 = ri.y * x00 + ri.x * x10;
            b11 = ri.y * x01 + ri.x * x11;
        }
        const int pi = rr_next(2 * I, m), qi = rr_next(2 * I + 1, m);
        const int pj = rr_next(2 * J, m), qj = rr_next(2 * J + 1, m);
        An[(long)pi * ne + pj] = b00;
        An[(long)pi * ne + qj] = b01;
        An[(long)qi * ne + pj] = b10;
        An[(long)qi * ne + qj] = b11;
        return;
    }
    const long ev = e - na;
    if (ev >= (long)m * n) return;
    const int I = (int)(ev / n), i = (int)(ev % n);
    const double3 r = cs[I];
    const double up = U[(long)(2 * I) * n + i], uq = U[(long)(2 * I + 1) * n + i];
    Un[(long)rr_next(2 * I, m) * n + i] = r.x * up - r.y * uq;
    Un[(long)rr_next(2 * I + 1, m) * n + i] = r.y * up + r.x * uq;
}

// Sum of squares of the off-diagonal and of all entries (n_e x n_e).
__global__ __launch_bounds__(256) void jac_norms(const double* __restrict__ A, int ne, double* __restrict__ out)
{
    __shared__ double so[256], sa[256];
    double o = 0.0, a = 0.0;
    for (long e = (long)blockIdx.x * 256 + threadIdx.x; e < (long)ne * ne; e += (long)gridDim.x * 256) {
        const double v = A[e] * A[e];
        a += v;
        if (e / ne != e % ne) o += v;
    }
    so[threadIdx.x] = o;
    sa[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            so[threadIdx.x] += so[threadIdx.x + w];
            sa[threadIdx.x] += sa[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(out, so[0]);
        atomicAdd(out + 1, sa[0]);
    }
}

// Rows of U in descending eigenvalue order (perm), each with its largest
// magnitude component (first on ties) made positive: V[k][i].
__global__ __launch_bounds__(256) void jac_sort_rows(const double* __restrict__ U, int n, const int* __restrict__ perm,
                                                     double* __restrict__ V)
{
    __shared__ double sv[256];
    __shared__ int si[256];
    const int k = blockIdx.x;
    const double* row = U + (long)perm[k] * n;
    double best = -1.0;
    int bi = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double a = fabs(row[i]);
        if (a > best) { best = a; bi = i; }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            const double b = sv[threadIdx.x + w];
            const int j = si[threadIdx.x + w];
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

// evecs[i][k] = V[k][i]
__global__ void jac_transpose(const double* __restrict__ V, int n, double* __restrict__ out)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e / n), k = (int)(e % n);
    out[e] = V[(long)k * n + i];
}

// PCA outputs: lambda[i][j] = V[i][j] / (norm_j sqrt(e_i)) for i < r,
// lambdainv[i][j] = norm_i V[j][i] sqrt(e_j) for j < r (V rows = eigenvectors).
__global__ void pca_outputs(const double* __restrict__ V, const double* __restrict__ ev,
                            const double* __restrict__ norm, int n, int r, double* __restrict__ lam,
                            double* __restrict__ linv)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)n * n) return;
    const int i = (int)(e / n), j = (int)(e % n);
    lam[e] = i < r ? V[(long)i * n + j] / (norm[j] * sqrt(ev[i])) : 0.0;
    linv[e] = j < r ? norm[i] * V[(long)j * n + i] * sqrt(ev[j]) : 0.0;
}

// ---------------------------------------------------------------------------
// Two-sided block Jacobi.  The n_e x n_e matrix (n_e a multiple of 2 BJ_B,
// padding rows and columns zero) is cut into nb = n_e / BJ_B blocks of BJ_B
// indices.  A step pairs the blocks disjointly (the circle method of a
// round-robin tournament over nb blocks: block nb-1 meets block s, block
// (s+k) mod (nb-1) meets (s-k) mod (nb-1)), so nb-1 steps meet every pair of
// blocks once (a sweep).  For pair k the 2 BJ_B x 2 BJ_B principal submatrix
// S_k = A[X_k, X_k] (X_k = the indices of its two blocks) is diagonalised in
// LDS by cyclic Jacobi, S_k = Q_k D_k Q_k^T, then
//   A <- P^T A P,  W <- P^T W,   P = the block-diagonal of the Q_k,
// block by block: A[X_a, X_b] <- Q_a^T A[X_a, X_b] Q_b (one workgroup per
// a <= b, writing both symmetric blocks in place) and W[X_k, cols] <-
// Q_k^T W[X_k, cols].  W's rows are the eigenvectors being accumulated.
// Padding indices meet only zero entries, so no rotation touches them.  The
// rotation and skip rules inside S_k are those of the scalar Jacobi above.
constexpr int BJ_B = 16;            // block size
constexpr int BJ_N = 2 * BJ_B;      // subproblem size
constexpr int BJ_INNER = 1;         // inner sweeps per subproblem (more do not cut the outer sweeps)

__host__ __device__ __forceinline__ void bj_pair(int s, int k, int nb, int& I, int& J)
{
    const int m = nb - 1;
    if (k == 0) {
        I = m;
        J = s;
    } else {
        I = (s + k) % m;
        J = (s - k + m) % m;
    }
}

// global index of local index a of pair (I, J)
__device__ __forceinline__ int bj_index(int a, int I, int J)
{
    return a < BJ_B ? I * BJ_B + a : J * BJ_B + (a - BJ_B);
}

// Diagonalise S_k for every pair k of step s: Qg[k] (BJ_N x BJ_N, row-major)
// with S_k Q = Q D; rotations applied counted into *nrot.
__global__ __launch_bounds__(128) void bj_solve(const double* __restrict__ A, int ne, int nb, int s,
                                                double abs_tol, double* __restrict__ Qg,
                                                unsigned long long* __restrict__ nrot)
{
    __shared__ double S[BJ_N][BJ_N + 1];
    __shared__ double Q[BJ_N][BJ_N + 1];
    __shared__ double rc[BJ_N / 2], rs[BJ_N / 2];
    __shared__ int rp[BJ_N / 2], rq[BJ_N / 2];
    __shared__ int changed;
    const int k = blockIdx.x, tid = threadIdx.x;
    int I, J;
    bj_pair(s, k, nb, I, J);
    for (int e = tid; e < BJ_N * BJ_N; e += 128) {
        const int a = e / BJ_N, b = e % BJ_N;
        S[a][b] = A[(long)bj_index(a, I, J) * ne + bj_index(b, I, J)];
        Q[a][b] = a == b ? 1.0 : 0.0;
    }
    unsigned long long applied = 0;
    if (tid == 0) changed = 0;
    __syncthreads();
    for (int sweep = 0; sweep < BJ_INNER; ++sweep) {
        for (int r = 0; r < BJ_N - 1; ++r) {
            if (tid < BJ_N / 2) {
                // pair tid of round r over BJ_N local indices (circle method)
                int p, q;
                if (tid == 0) {
                    p = BJ_N - 1;
                    q = r;
                } else {
                    p = (r + tid) % (BJ_N - 1);
                    q = (r - tid + BJ_N - 1) % (BJ_N - 1);
                }
                const double apq = S[p][q], app = S[p][p], aqq = S[q][q];
                double c = 1.0, sn = 0.0;
                if (fabs(apq) > fmax(1e-16 * sqrt(fabs(app * aqq)), abs_tol)) {
                    const double tau = (aqq - app) / (2.0 * apq);
                    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                    c = 1.0 / sqrt(1.0 + t * t);
                    sn = t * c;
                    changed = 1;
                    ++applied;
                }
                rp[tid] = p;
                rq[tid] = q;
                rc[tid] = c;
                rs[tid] = sn;
            }
            __syncthreads();
            // S <- J^T S J in 2 x 2 blocks (pair u rows, pair v columns); Q <- Q J
            for (int e = tid; e < (BJ_N / 2) * (BJ_N / 2); e += 128) {
                const int u = e / (BJ_N / 2), v = e % (BJ_N / 2);
                const int p = rp[u], q = rq[u], x = rp[v], y = rq[v];
                const double cu = rc[u], su = rs[u], cv = rc[v], sv = rs[v];
                const double s_px = S[p][x], s_py = S[p][y], s_qx = S[q][x], s_qy = S[q][y];
                // rows: p' = c p - s q, q' = s p + c q
                const double t_px = cu * s_px - su * s_qx, t_py = cu * s_py - su * s_qy;
                const double t_qx = su * s_px + cu * s_qx, t_qy = su * s_py + cu * s_qy;
                // columns: x' = c x - s y, y' = s x + c y
                S[p][x] = cv * t_px - sv * t_py;
                S[p][y] = sv * t_px + cv * t_py;
                S[q][x] = cv * t_qx - sv * t_qy;
                S[q][y] = sv * t_qx + cv * t_qy;
            }
            for (int e = tid; e < (BJ_N / 2) * BJ_N; e += 128) {
                const int v = e / BJ_N, row = e % BJ_N;
                const int x = rp[v], y = rq[v];
                const double cv = rc[v], sv = rs[v];
                const double qx = Q[row][x], qy = Q[row][y];
                Q[row][x] = cv * qx - sv * qy;
                Q[row][y] = sv * qx + cv * qy;
            }
            __syncthreads();
        }
        const int ch = changed;   // read by every thread before the reset below
        __syncthreads();
        if (!ch) break;
        if (tid == 0) changed = 0;
        __syncthreads();
    }
    // One Newton-Schulz step, Q <- Q (3 I - Q^T Q) / 2, takes the product of
    // up to BJ_INNER (BJ_N - 1) rotations back to orthogonality (the error
    // goes from k eps to (k eps)^2), so that repeated block updates do not
    // scale A or W.  S is free after the sweeps.
    __syncthreads();
    for (int e = tid; e < BJ_N * BJ_N; e += 128) {
        const int i = e / BJ_N, j = e % BJ_N;
        double m = 0.0;
        for (int l = 0; l < BJ_N; ++l) m = fma(Q[l][i], Q[l][j], m);
        S[i][j] = (i == j ? 1.5 : 0.0) - 0.5 * m;
    }
    __syncthreads();
    for (int e = tid; e < BJ_N * BJ_N; e += 128) {
        const int a = e / BJ_N, b = e % BJ_N;
        double q = 0.0;
        for (int l = 0; l < BJ_N; ++l) q = fma(Q[a][l], S[l][b], q);
        Qg[(long)k * BJ_N * BJ_N + e] = q;
    }
    if (applied) atomicAdd(nrot, applied);
}

// A[X_a, X_b] <- Q_a^T A[X_a, X_b] Q_b for pairs a <= b of step s, one wave
// per block (both symmetric blocks written; in place, each block belongs to
// one workgroup).  FP64 MFMA 16x16x4: T = Q_a^T A_ab, then R = T Q_b, each a
// 2 x 2 grid of 16 x 16 blocks over K = 32; R goes through LDS so that both
// stores run along rows.
__global__ __launch_bounds__(64) void bj_update_A(double* __restrict__ A, int ne, int nb, int s,
                                                  const double* __restrict__ Qg)
{
    static_assert(BJ_N == 32, "one wave per 32 x 32 block");
    __shared__ double T[BJ_N][BJ_N + 1];
    const int P = nb / 2;
    int idx = blockIdx.x, a = 0;
    while (idx >= P - a) {
        idx -= P - a;
        ++a;
    }
    const int b = a + idx;
    int Ia, Ja, Ib, Jb;
    bj_pair(s, a, nb, Ia, Ja);
    bj_pair(s, b, nb, Ib, Jb);
    const int lane = threadIdx.x, lr = lane & 15, lk = lane >> 4;
    // the rotations are read as MFMA operands straight from Qg (cache-resident)
    const double* const Qa = Qg + (long)a * BJ_N * BJ_N;
    const double* const Qb = Qg + (long)b * BJ_N * BJ_N;
    // B operand of T: A_ab[l][j], l = 4 ks + lane/16, j = 16 c + lane%16
    double bv[BJ_N / 4][2];
#pragma unroll
    for (int ks = 0; ks < BJ_N / 4; ++ks) {
        const long row = (long)bj_index(4 * ks + lk, Ia, Ja) * ne;
#pragma unroll
        for (int c = 0; c < 2; ++c) bv[ks][c] = A[row + bj_index(16 * c + lr, Ib, Jb)];
    }
    __syncthreads();
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[x][c] = (fd4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < BJ_N / 4; ++ks)
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const double av = Qa[(4 * ks + lk) * BJ_N + 16 * x + lr];   // (Q_a^T)[i][l] = Q_a[l][i]
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[x][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[ks][c], acc[x][c], 0, 0, 0);
        }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) T[16 * x + 4 * v + lk][16 * c + lr] = acc[x][c][v];
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[x][c] = (fd4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < BJ_N / 4; ++ks)
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const double av = T[16 * x + lr][4 * ks + lk];
#pragma unroll
            for (int c = 0; c < 2; ++c)
                acc[x][c] =
                    __builtin_amdgcn_mfma_f64_16x16x4f64(av, Qb[(4 * ks + lk) * BJ_N + 16 * c + lr], acc[x][c], 0, 0, 0);
        }
    __syncthreads();
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) T[16 * x + 4 * v + lk][16 * c + lr] = acc[x][c][v];
    __syncthreads();
    // R[i][j] to A[X_a[i]][X_b[j]] and A[X_b[j]][X_a[i]], rows contiguous
    for (int e = lane; e < BJ_N * BJ_N; e += 64) {
        const int i = e / BJ_N, j = e % BJ_N;
        A[(long)bj_index(i, Ia, Ja) * ne + bj_index(j, Ib, Jb)] = T[i][j];
        if (a != b) A[(long)bj_index(i, Ib, Jb) * ne + bj_index(j, Ia, Ja)] = T[j][i];
    }
}

// W[X_k, c0 .. c0 + 256) <- Q_k^T W[X_k, c0 .. c0 + 256) for every pair k of
// step s (W: n_e rows of n, row-major), FP64 MFMA: wave w takes 64 columns,
// its B operands read from W straight into registers before any store.
__global__ __launch_bounds__(256) void bj_update_W(double* __restrict__ W, int n, int nb, int s,
                                                   const double* __restrict__ Qg)
{
    static_assert(BJ_N == 32, "2 x 4 blocks of 16 x 16 per wave");
    const int ncb = (n + 255) / 256;
    const int k = blockIdx.x / ncb, c0 = (blockIdx.x % ncb) * 256;
    int I, J;
    bj_pair(s, k, nb, I, J);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, lr = lane & 15, lk = lane >> 4;
    const double* const Qk = Qg + (long)k * BJ_N * BJ_N;   // read as MFMA operands from cache
    const int cb = c0 + 64 * wave + lr;
    double bv[BJ_N / 4][4];
#pragma unroll
    for (int ks = 0; ks < BJ_N / 4; ++ks) {
        const long row = (long)bj_index(4 * ks + lk, I, J) * n;
#pragma unroll
        for (int c = 0; c < 4; ++c) bv[ks][c] = cb + 16 * c < n ? W[row + cb + 16 * c] : 0.0;
    }
    fd4 acc[2][4];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[x][c] = (fd4){0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int ks = 0; ks < BJ_N / 4; ++ks)
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const double av = Qk[(4 * ks + lk) * BJ_N + 16 * x + lr];
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[x][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv[ks][c], acc[x][c], 0, 0, 0);
        }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long row = (long)bj_index(16 * x + 4 * v + lk, I, J) * n;
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (cb + 16 * c < n) W[row + cb + 16 * c] = acc[x][c][v];
        }
}

// Jacobi on the device.  On return d_V (n x n) holds the sorted, sign-fixed
// eigenvectors as rows and evals (host) the eigenvalues, descending.
static int jacobi_device(mcmb_ctx* c, int n, const double* d_G, int normalise, double* d_norm, double* evals,
                         DevBuf& d_V, int* sweeps_out)
{
    const int ne = n + (n & 1), m = ne / 2;
    DevBuf A0, A1, U0, U1, dcs, dcnt, dnrm, dperm;
    MCMB_TRY(A0.alloc(sizeof(double) * (size_t)ne * ne));
    MCMB_TRY(A1.alloc(sizeof(double) * (size_t)ne * ne));
    MCMB_TRY(U0.alloc(sizeof(double) * (size_t)ne * n));
    MCMB_TRY(U1.alloc(sizeof(double) * (size_t)ne * n));
    MCMB_TRY(dcs.alloc(sizeof(double3) * (size_t)m));
    MCMB_TRY(dcnt.alloc(sizeof(unsigned long long)));
    MCMB_TRY(dnrm.alloc(sizeof(double) * 2));
    const long ninit = std::max<long>((long)ne * ne, (long)ne * n);
    MCMB_TRY(launch_guard("Jacobi init", (ninit + 255) / 256, 256));
    hipLaunchKernelGGL(jac_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, c->stream, d_G, n, ne,
                       normalise, d_norm, A0.as<double>(), U0.as<double>());
    MCMB_HIP(hipGetLastError());
    const long napply = (long)m * m + (long)m * n;
    MCMB_TRY(launch_guard("Jacobi step", (napply + 255) / 256, 256));
    double* A = A0.as<double>();
    double* An = A1.as<double>();
    double* U = U0.as<double>();
    double* Un = U1.as<double>();
    int sweep = 0;
    bool converged = false;
    double abs_tol = 0.0;
    for (; sweep < 50 && !converged; ++sweep) {
        MCMB_HIP(hipMemsetAsync(dnrm.p, 0, sizeof(double) * 2, c->stream));
        hipLaunchKernelGGL(jac_norms, dim3(1024), dim3(256), 0, c->stream, A, ne, dnrm.as<double>());
        MCMB_HIP(hipGetLastError());
        double nr[2];
        MCMB_HIP(hipMemcpyAsync(nr, dnrm.p, sizeof(nr), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        const double stop = std::max(1e-15, 2e-17 * n);
        if (nr[0] <= stop * stop * nr[1]) { converged = true; break; }
        if (sweep == 0) abs_tol = 1e-17 * sqrt(nr[1]);
        MCMB_HIP(hipMemsetAsync(dcnt.p, 0, sizeof(unsigned long long), c->stream));
        for (int step = 0; step < ne - 1; ++step) {
            hipLaunchKernelGGL(jac_rot, dim3((m + 255) / 256), dim3(256), 0, c->stream, A, ne, abs_tol,
                               dcs.as<double3>(), dcnt.as<unsigned long long>());
            hipLaunchKernelGGL(jac_apply, dim3((unsigned)((napply + 255) / 256)), dim3(256), 0, c->stream, A, An,
                               U, Un, dcs.as<double3>(), ne, n);
            std::swap(A, An);
            std::swap(U, Un);
        }
        MCMB_HIP(hipGetLastError());
        unsigned long long cnt = 0;
        MCMB_HIP(hipMemcpyAsync(&cnt, dcnt.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        if (cnt == 0) { converged = true; ++sweep; break; }
    }
    if (!converged) return fail(MCMB_ERR_ARG, "Jacobi eigen-decomposition of n=%d did not converge in 50 sweeps", n);
    // layout is back in index order: eigenvalue of index i = A[i][i]
    std::vector<double> diag(n);
    MCMB_HIP(hipMemcpy2DAsync(diag.data(), sizeof(double), A, sizeof(double) * (ne + 1), sizeof(double), n,
                              hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return diag[a] > diag[b]; });
    for (int k = 0; k < n; ++k) evals[k] = diag[perm[k]];
    MCMB_TRY(dperm.alloc(sizeof(int) * (size_t)n));
    MCMB_HIP(hipMemcpyAsync(dperm.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(d_V.alloc(sizeof(double) * (size_t)n * n));
    hipLaunchKernelGGL(jac_sort_rows, dim3(n), dim3(256), 0, c->stream, U, n, dperm.as<int>(), d_V.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));   // perm and the work buffers go out of scope
    if (sweeps_out) *sweeps_out = sweep;
    return MCMB_OK;
}

// Block Jacobi on the device; same outputs as jacobi_device.
static int block_jacobi_device(mcmb_ctx* c, int n, const double* d_G, int normalise, double* d_norm, double* evals,
                               DevBuf& d_V, int* sweeps_out)
{
    const int ne = (n + BJ_N - 1) / BJ_N * BJ_N, nb = ne / BJ_B, P = nb / 2;
    const int ncb = (n + 255) / 256;
    DevBuf A, W, Qg, dcnt, dnrm, dperm;
    MCMB_TRY(A.alloc(sizeof(double) * (size_t)ne * ne));
    MCMB_TRY(W.alloc(sizeof(double) * (size_t)ne * n));
    // two sets of rotations: step s writes set s % 2 (see the W stream below)
    const size_t qset = (size_t)P * BJ_N * BJ_N;
    MCMB_TRY(Qg.alloc(sizeof(double) * 2 * qset));
    // The eigenvector rows W are rotated on a second stream (default
    // priority): W(s) needs only step s's rotations, and the next steps'
    // solves and A updates need only A, so W(s) overlaps solve(s + 1) and
    // A(s + 1).  Step s + 2 reuses the rotation set of step s after W(s).
    struct Side {
        hipStream_t st = nullptr;
        hipEvent_t solved[2] = {nullptr, nullptr}, rotated[2] = {nullptr, nullptr};
        // release on every exit path; a failure here has no caller to report to
        ~Side()
        {
            if (st) (void)hipStreamSynchronize(st);
            for (int k = 0; k < 2; ++k) {
                if (solved[k]) (void)hipEventDestroy(solved[k]);
                if (rotated[k]) (void)hipEventDestroy(rotated[k]);
            }
            if (st) (void)hipStreamDestroy(st);
        }
    } side;
    MCMB_HIP(hipStreamCreate(&side.st));
    for (int k = 0; k < 2; ++k) {
        MCMB_HIP(hipEventCreateWithFlags(&side.solved[k], hipEventDisableTiming));
        MCMB_HIP(hipEventCreateWithFlags(&side.rotated[k], hipEventDisableTiming));
    }
    long gstep = 0;   // steps issued so far, over all sweeps
    MCMB_TRY(dcnt.alloc(sizeof(unsigned long long)));
    MCMB_TRY(dnrm.alloc(sizeof(double) * 2));
    const long ninit = std::max<long>((long)ne * ne, (long)ne * n);
    MCMB_TRY(launch_guard("block Jacobi init", (ninit + 255) / 256, 256));
    hipLaunchKernelGGL(jac_init, dim3((unsigned)((ninit + 255) / 256)), dim3(256), 0, c->stream, d_G, n, ne,
                       normalise, d_norm, A.as<double>(), W.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_TRY(launch_guard("block Jacobi A update", (unsigned long long)P * (P + 1) / 2, 64));
    MCMB_TRY(launch_guard("block Jacobi W update", (unsigned long long)P * ncb, 256));
    int sweep = 0;
    bool converged = false;
    double abs_tol = 0.0;
    for (; sweep < 60 && !converged; ++sweep) {
        MCMB_HIP(hipMemsetAsync(dnrm.p, 0, sizeof(double) * 2, c->stream));
        hipLaunchKernelGGL(jac_norms, dim3(1024), dim3(256), 0, c->stream, A.as<double>(), ne, dnrm.as<double>());
        MCMB_HIP(hipGetLastError());
        double nr[2];
        MCMB_HIP(hipMemcpyAsync(nr, dnrm.p, sizeof(nr), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        const double stop = std::max(1e-15, 2e-17 * n);
        if (nr[0] <= stop * stop * nr[1]) { converged = true; break; }
        if (sweep == 0) abs_tol = 1e-17 * sqrt(nr[1]);
        MCMB_HIP(hipMemsetAsync(dcnt.p, 0, sizeof(unsigned long long), c->stream));
        for (int step = 0; step < nb - 1; ++step, ++gstep) {
            const int q = (int)(gstep & 1);
            double* const Qs = Qg.as<double>() + q * qset;
            if (gstep >= 2) MCMB_HIP(hipStreamWaitEvent(c->stream, side.rotated[q], 0));
            hipLaunchKernelGGL(bj_solve, dim3(P), dim3(128), 0, c->stream, (const double*)A.as<double>(), ne, nb, step,
                               abs_tol, Qs, dcnt.as<unsigned long long>());
            MCMB_HIP(hipEventRecord(side.solved[q], c->stream));
            hipLaunchKernelGGL(bj_update_A, dim3((unsigned)((long)P * (P + 1) / 2)), dim3(64), 0, c->stream,
                               A.as<double>(), ne, nb, step, (const double*)Qs);
            MCMB_HIP(hipStreamWaitEvent(side.st, side.solved[q], 0));
            hipLaunchKernelGGL(bj_update_W, dim3((unsigned)((long)P * ncb)), dim3(256), 0, side.st, W.as<double>(), n,
                               nb, step, (const double*)Qs);
            MCMB_HIP(hipEventRecord(side.rotated[q], side.st));
        }
        MCMB_HIP(hipGetLastError());
        unsigned long long cnt = 0;
        MCMB_HIP(hipMemcpyAsync(&cnt, dcnt.p, sizeof(cnt), hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        if (cnt == 0) { converged = true; ++sweep; break; }
    }
    // W complete before it is read
    if (gstep > 0) MCMB_HIP(hipStreamWaitEvent(c->stream, side.rotated[(gstep - 1) & 1], 0));
    if (!converged)
        return fail(MCMB_ERR_ARG, "block Jacobi eigen-decomposition of n=%d did not converge in 60 sweeps", n);
    std::vector<double> diag(n);
    MCMB_HIP(hipMemcpy2DAsync(diag.data(), sizeof(double), A.as<double>(), sizeof(double) * (ne + 1), sizeof(double),
                              n, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    std::vector<int> perm(n);
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return diag[a] > diag[b]; });
    for (int k = 0; k < n; ++k) evals[k] = diag[perm[k]];
    MCMB_TRY(dperm.alloc(sizeof(int) * (size_t)n));
    MCMB_HIP(hipMemcpyAsync(dperm.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(d_V.alloc(sizeof(double) * (size_t)n * n));
    hipLaunchKernelGGL(jac_sort_rows, dim3(n), dim3(256), 0, c->stream, (const double*)W.as<double>(), n,
                       dperm.as<int>(), d_V.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipStreamSynchronize(c->stream));   // perm and the work buffers go out of scope
    if (sweeps_out) *sweeps_out = sweep;
    return MCMB_OK;
}

// Eigen-decomposition: Householder tridiagonalisation (sweeps 0) up to its
// size limit, else block Jacobi; MCMB_EIGEN=block or scalar selects the block
// or the scalar Jacobi.
static int eigen_device(mcmb_ctx* c, int n, const double* d_G, int normalise, double* d_norm, double* evals,
                        DevBuf& d_V, int* sweeps_out)
{
    const char* e = getenv("MCMB_EIGEN");
    if (e && !strcmp(e, "scalar")) return jacobi_device(c, n, d_G, normalise, d_norm, evals, d_V, sweeps_out);
    if ((e && !strcmp(e, "block")) || !tridiag_supported(n))
        return block_jacobi_device(c, n, d_G, normalise, d_norm, evals, d_V, sweeps_out);
    if (sweeps_out) *sweeps_out = 0;
    return tridiag_eigen(c, n, d_G, normalise, d_norm, evals, d_V);
}

}  // namespace mcmb

using namespace mcmb;

extern "C" int mcmb_eigen_sym(mcmb_ctx* c, int n, const double* a, double* evals, double* evecs, int* sweeps)
{
    if (!c || !a || !evals || n < 1) return fail(MCMB_ERR_ARG, "mcmb_eigen_sym: bad arguments");
    MCMB_TRY(set_device(c));
    DevBuf dG, dV, dE;
    MCMB_TRY(dG.alloc(sizeof(double) * (size_t)n * n));
    MCMB_TRY(copy_to_device(c, dG.p, a, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    MCMB_TRY(eigen_device(c, n, dG.as<double>(), 0, nullptr, evals, dV, sweeps));
    if (evecs) {
        MCMB_TRY(dE.alloc(sizeof(double) * (size_t)n * n));
        const long nn = (long)n * n;
        hipLaunchKernelGGL(jac_transpose, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream,
                           dV.as<double>(), n, dE.as<double>());
        MCMB_HIP(hipGetLastError());
    }
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    if (evecs) MCMB_TRY(copy_to_host(c, evecs, dE.p, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 0;
    return MCMB_OK;
}

extern "C" int mcmb_orthonormalise_late(mcmb_ctx* c, int n, const double* gamma, int pca_flag, double* lambda,
                                        double* lambdainv, double* norm, double* evals, int* rank)
{
    if (!c || !gamma || !lambda || !lambdainv || !norm || !evals || n < 1)
        return fail(MCMB_ERR_ARG, "mcmb_orthonormalise_late: bad arguments");
    MCMB_TRY(set_device(c));
    for (int i = 0; i < n; ++i)
        if (!(gamma[(long)i * n + i] > 0.0))
            return fail(MCMB_ERR_ARG, "mcmb_orthonormalise_late: Gamma[%d][%d] = %g is not positive", i, i,
                        gamma[(long)i * n + i]);
    DevBuf dG, dN, dV, dEv, dL, dLi;
    const size_t nn = (size_t)n * n;
    MCMB_TRY(dG.alloc(sizeof(double) * nn));
    MCMB_TRY(dN.alloc(sizeof(double) * n));
    MCMB_TRY(copy_to_device(c, dG.p, gamma, sizeof(double) * nn));
    MCMB_TRY(eigen_device(c, n, dG.as<double>(), 1, dN.as<double>(), evals, dV, nullptr));
    MCMB_HIP(hipMemcpyAsync(norm, dN.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    auto pca = [&]() -> int {
        int r = 0;
        while (r < n && evals[r] > 0.0) ++r;
        if (r == 0) return fail(MCMB_ERR_NOT_SPD, "mcmb_orthonormalise_late: no positive eigenvalue");
        MCMB_TRY(dEv.alloc(sizeof(double) * n));
        MCMB_TRY(dL.alloc(sizeof(double) * nn));
        MCMB_TRY(dLi.alloc(sizeof(double) * nn));
        MCMB_HIP(hipMemcpyAsync(dEv.p, evals, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(pca_outputs, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream,
                           dV.as<double>(), dEv.as<double>(), dN.as<double>(), n, r, dL.as<double>(),
                           dLi.as<double>());
        MCMB_HIP(hipGetLastError());
        MCMB_TRY(copy_to_host(c, lambda, dL.p, sizeof(double) * nn));
        MCMB_TRY(copy_to_host(c, lambdainv, dLi.p, sizeof(double) * nn));
        if (rank) *rank = r;
        return MCMB_OK;
    };
    if (pca_flag == 0 || evals[n - 1] < 0.0) return pca();
    // Cholesky branch: lambdainv = D L, lambda = (D L)^-1, lower triangles.  A
    // Gram whose eigenvalues all come out >= 0 can still fail the FP64
    // Cholesky (its smallest ones are rounding noise); the reference would
    // abort there, this takes the PCA branch instead.
    std::vector<double> ortho(nn), lam(nn);
    int bad = -1;
    const int st = mcmb_orthonormalise(c, n, gamma, ortho.data(), lam.data(), norm, &bad);
    if (st == MCMB_ERR_NOT_SPD) {
        MCMB_HIP(hipMemcpyAsync(norm, dN.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        return pca();
    }
    MCMB_TRY(st);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            lambdainv[(long)i * n + j] = j <= i ? ortho[(long)i * n + j] : 0.0;
            lambda[(long)i * n + j] = j <= i ? lam[(long)i * n + j] : 0.0;
        }
    if (rank) *rank = n;
    return MCMB_OK;
}
