// FP64 orthonormalisation on the GPU (ortho_prim_mpi.c:198-284) and the
// Cholesky solve of the projection (eigenmode_mpi.c:195-219).
//
//   norm_i = sqrt(G_ii),  C = D^-1 G D^-1 = L L^T  (blocked right-looking
//   Cholesky, block NB: one-workgroup LDS factorisation of the diagonal block,
//   row-parallel triangular solve of the panel, SYRK update of the trailing
//   lower triangle on FP64 MFMA), Linv = L^-1 by blocked right-looking
//   forward substitution on the identity (each block's update of the rows
//   below it on FP64 MFMA, over the right-hand sides that are not yet zero).
//   ortho  = symmetrised D L          (update_ortho, fixed_data.c:3285-3295)
//   lambda = symmetrised (D L)^-1     (update_lambda, fixed_data.c:3576-3586)
// A non-positive pivot sets a device flag (first failing column + 1); later
// kernels see it and return, and the host reports MCMB_ERR_NOT_SPD.
#include "mcmb_internal.h"
#include <algorithm>
#include <vector>

namespace mcmb {

constexpr int NB = 32;     // Cholesky / substitution block

__global__ void la_norm_kernel(const double* __restrict__ G, int n, double* __restrict__ norm)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) norm[i] = sqrt(G[(size_t)i * n + i]);
}

__global__ void la_scale_kernel(const double* __restrict__ G, int n, const double* __restrict__ norm,
                                double* __restrict__ C)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int i = (int)(e / n), j = (int)(e % n);
    C[e] = G[e] / (norm[i] * norm[j]);
}

// Factor the nb x nb diagonal block at (k0, k0) in LDS (unblocked, right-looking).
__global__ __launch_bounds__(256) void la_potrf_diag(double* __restrict__ C, int n, int k0, int* __restrict__ flag)
{
    __shared__ double T[NB][NB + 1];
    __shared__ int bad;
    if (*flag) return;
    const int nb = min(NB, n - k0);
    const int tid = threadIdx.x;
    for (int e = tid; e < NB * NB; e += 256) {
        int i = e / NB, j = e % NB;
        T[i][j] = (i < nb && j < nb) ? C[(size_t)(k0 + i) * n + k0 + j] : 0.0;
    }
    if (tid == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < nb; ++j) {
        if (tid == 0) {
            double d = T[j][j];
            if (!(d > 0.0)) bad = k0 + j + 1;
            else T[j][j] = sqrt(d);
        }
        __syncthreads();
        if (bad) break;
        if (tid > j && tid < nb) T[tid][j] /= T[j][j];
        __syncthreads();
        for (int e = tid; e < NB * NB; e += 256) {
            int i = e / NB, l = e % NB;
            if (l > j && i >= l && i < nb) T[i][l] -= T[i][j] * T[l][j];
        }
        __syncthreads();
    }
    if (bad) {
        if (tid == 0) atomicCAS(flag, 0, bad);
        return;
    }
    for (int e = tid; e < nb * nb; e += 256) {
        int i = e / nb, j = e % nb;
        if (j <= i) C[(size_t)(k0 + i) * n + k0 + j] = T[i][j];
    }
}

// Panel rows i >= k0+nb:  L[i, k0:k0+nb] = C[i, k0:k0+nb] L_kk^-T (one row per thread).
__global__ __launch_bounds__(256) void la_trsm_panel(double* __restrict__ C, int n, int k0, const int* __restrict__ flag)
{
    __shared__ double Lk[NB][NB + 1];
    if (*flag) return;
    const int nb = min(NB, n - k0);
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        int i = e / NB, j = e % NB;
        Lk[i][j] = (i < nb && j <= i) ? C[(size_t)(k0 + i) * n + k0 + j] : 0.0;
    }
    __syncthreads();
    const int row = k0 + nb + blockIdx.x * 256 + threadIdx.x;
    if (row >= n) return;
    double x[NB];
    double* cr = C + (size_t)row * n + k0;
#pragma unroll
    for (int j = 0; j < NB; ++j) x[j] = j < nb ? cr[j] : 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
        double s = x[j];
#pragma unroll
        for (int l = 0; l < j; ++l) s -= x[l] * Lk[j][l];
        x[j] = (j < nb) ? s / Lk[j][j] : 0.0;
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
        if (j < nb) cr[j] = x[j];
}

typedef double fd4 __attribute__((ext_vector_type(4)));

// C(i, j) -= sum_{k < K} A(i, k) B(j, k) over 64 x 64 tiles of rows [i0, i1)
// and columns [j0, j1), on FP64 MFMA (v_mfma_f64_16x16x4f64), with
// A(i, k) = A[i ai + k ak], B(j, k) = B[j bj + k bk], C(i, j) = C[i ci + j cj].
// LOWER: only the tiles I >= J of a square range (i0 = j0, i1 = j1), and
// only j <= i written (the SYRK of the Cholesky).  Four waves per tile, a
// 32 x 32 quadrant each (2 x 2 blocks of 16 x 16), operands straight from
// global memory, K in steps of 4.  Returns at once when *flag is set.
template <bool LOWER>
__global__ __launch_bounds__(256) void la_mfma_update(const double* __restrict__ A, long ai, long ak,
                                                      const double* __restrict__ B, long bj, long bk,
                                                      double* __restrict__ C, long ci, long cj, int i0, int i1,
                                                      int j0, int j1, int K, const int* __restrict__ flag)
{
    if (flag && *flag) return;
    int I, J;
    if (LOWER) {
        // lower tile index t -> (I, J), I >= J
        const int t = blockIdx.x;
        I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
        while ((I + 1) * (I + 2) / 2 <= t) ++I;
        while (I * (I + 1) / 2 > t) --I;
        J = t - I * (I + 1) / 2;
    } else {
        const int ntj = (j1 - j0 + 63) / 64;
        I = blockIdx.x / ntj;
        J = blockIdx.x % ntj;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int rb = i0 + 64 * I + 32 * (w >> 1), cb = j0 + 64 * J + 32 * (w & 1);
    const int l16 = lane & 15, kq = lane >> 4;
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (fd4){0.0, 0.0, 0.0, 0.0};
    const int r0 = rb + l16, r1 = rb + 16 + l16, c0 = cb + l16, c1 = cb + 16 + l16;
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + kq;
        const bool kin = k < K;
        double a[2], b[2];
        a[0] = (kin && r0 < i1) ? A[r0 * ai + k * ak] : 0.0;
        a[1] = (kin && r1 < i1) ? A[r1 * ai + k * ak] : 0.0;
        b[0] = (kin && c0 < j1) ? B[c0 * bj + k * bk] : 0.0;
        b[1] = (kin && c1 < j1) ? B[c1 * bj + k * bk] : 0.0;
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
            for (int y = 0; y < 2; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
    // element v of a block: row 4 v + lane / 16, column lane % 16
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rb + 16 * x + 4 * v + kq, j = cb + 16 * y + l16;
                if (i < i1 && j < j1 && (!LOWER || j <= i)) C[i * ci + j * cj] -= acc[x][y][v];
            }
}

// Diagonal-block substitution on every RHS (one RHS per thread):
// lower (upper == 0): X[rb+j] = (X[rb+j] - sum_{l<j} L(rb+j, rb+l) X[rb+l]) / L(rb+j, rb+j)
// upper (upper == 1, U = L^T): backward over j.  L is row-major lower.
__global__ __launch_bounds__(256) void la_trsm_rhs(const double* __restrict__ Lm, int n, int upper,
                                                   double* __restrict__ X, int ldx, int nrhs, int rb)
{
    __shared__ double Lk[NB][NB + 1];
    const int nb = min(NB, n - rb);
    for (int e = threadIdx.x; e < NB * NB; e += 256) {
        int i = e / NB, j = e % NB;
        Lk[i][j] = (i < nb && j <= i) ? Lm[(size_t)(rb + i) * n + rb + j] : 0.0;
    }
    __syncthreads();
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= nrhs) return;
    double* x = X + (size_t)r * ldx + rb;
    double v[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) v[j] = j < nb ? x[j] : 0.0;
    if (!upper) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
            double s = v[j];
#pragma unroll
            for (int l = 0; l < j; ++l) s -= Lk[j][l] * v[l];
            v[j] = j < nb ? s / Lk[j][j] : 0.0;
        }
    } else {
#pragma unroll
        for (int j = NB - 1; j >= 0; --j) {
            double s = v[j];
#pragma unroll
            for (int l = j + 1; l < NB; ++l) s -= Lk[l][j] * v[l];
            v[j] = j < nb ? s / Lk[j][j] : 0.0;
        }
    }
#pragma unroll
    for (int j = 0; j < NB; ++j)
        if (j < nb) x[j] = v[j];
}

__global__ void la_identity(double* __restrict__ X, int n)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (size_t)n * n) X[e] = ((int)(e / n) == (int)(e % n)) ? 1.0 : 0.0;
}

// ortho[i][j] = ortho[j][i] = norm[i] L[i][j] (i >= j);
// lambda[i][j] = lambda[j][i] = Linv[i][j] / norm[j], Linv RHS-major (Linv[c*n + i] = Linv(i, c)).
__global__ void la_outputs(const double* __restrict__ L, const double* __restrict__ Xinv,
                           const double* __restrict__ norm, int n, double* __restrict__ ortho,
                           double* __restrict__ lambda)
{
    size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    int i = (int)(e / n), j = (int)(e % n);
    int a = i > j ? i : j, b = i > j ? j : i;
    ortho[e] = norm[a] * L[(size_t)a * n + b];
    lambda[e] = Xinv[(size_t)b * n + a] / norm[b];
}

// Forward (L y = b) or backward (L^T y = b) substitution on nrhs RHS-major
// columns (X[r n + i]), right-looking: per block, the diagonal block's
// substitution, then the update of the rows after it (forward) or before it
// (backward) on FP64 MFMA.  inverse (forward only, X = I on entry): right-hand
// side r is zero in the rows of the blocks before its own, so block b works on
// the right-hand sides r < rb + nb only.
static int la_solve(mcmb_ctx* c, const double* dL, int n, int upper, double* dX, int nrhs, bool inverse = false)
{
    const int nblk = (n + NB - 1) / NB;
    for (int b = 0; b < nblk; ++b) {
        const int kb = upper ? (nblk - 1 - b) : b;
        const int rb = kb * NB, nb = std::min(NB, n - rb);
        const int nr = inverse ? std::min(nrhs, rb + nb) : nrhs;
        hipLaunchKernelGGL(la_trsm_rhs, dim3((unsigned)((nr + 255) / 256)), dim3(256), 0, c->stream, dL, n, upper, dX,
                           n, nr, rb);
        // rows [i0, i1) of the update; columns = the right-hand sides [0, nr)
        const int i0 = upper ? 0 : rb + nb, i1 = upper ? rb : n;
        if (i1 > i0) {
            const unsigned tiles = (unsigned)(((i1 - i0 + 63) / 64) * ((nr + 63) / 64));
            // forward: A(i, k) = L(i, rb + k); backward: A(i, k) = L^T(i, rb + k) = L(rb + k, i)
            const double* A = upper ? dL + (size_t)rb * n : dL + rb;
            hipLaunchKernelGGL((la_mfma_update<false>), dim3(tiles), dim3(256), 0, c->stream, A,
                               upper ? 1l : (long)n, upper ? (long)n : 1l, (const double*)(dX + rb), (long)n, 1l, dX,
                               1l, (long)n, i0, i1, 0, nr, nb, (const int*)nullptr);
        }
    }
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// In-place Cholesky (lower) of the n x n matrix at dC; *bad = failing column+1.
static int la_potrf(mcmb_ctx* c, double* dC, int n, int* d_flag, int* bad)
{
    MCMB_HIP(hipMemsetAsync(d_flag, 0, sizeof(int), c->stream));
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int k1 = k0 + NB < n ? k0 + NB : n;
        hipLaunchKernelGGL(la_potrf_diag, dim3(1), dim3(256), 0, c->stream, dC, n, k0, d_flag);
        if (k1 < n) {
            hipLaunchKernelGGL(la_trsm_panel, dim3((unsigned)((n - k1 + 255) / 256)), dim3(256), 0, c->stream, dC, n, k0,
                               d_flag);
            // C(i, j) -= sum_l L(i, k0 + l) L(j, k0 + l) on the trailing lower triangle
            const int T = (n - k1 + 63) / 64;
            hipLaunchKernelGGL((la_mfma_update<true>), dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, c->stream,
                               (const double*)(dC + k0), (long)n, 1l, (const double*)(dC + k0), (long)n, 1l, dC, (long)n,
                               1l, k1, n, k1, n, k1 - k0, (const int*)d_flag);
        }
    }
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipMemcpyAsync(bad, d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

}  // namespace mcmb

using namespace mcmb;

extern "C" int mcmb_orthonormalise(mcmb_ctx* c, int n, const double* gamma, double* ortho, double* lambda,
                                   double* norm, int* bad_pivot)
{
    if (!c || !gamma || n < 1) return fail(MCMB_ERR_ARG, "mcmb_orthonormalise: bad argument");
    MCMB_TRY(set_device(c));
    const size_t nn = (size_t)n * n;
    DevBuf dG, dC, dX, dN, dF, dO, dL;
    MCMB_TRY(dG.alloc(nn * sizeof(double)));
    MCMB_TRY(dC.alloc(nn * sizeof(double)));
    MCMB_TRY(dN.alloc(n * sizeof(double)));
    MCMB_TRY(dF.alloc(sizeof(int)));
    MCMB_TRY(copy_to_device(c, dG.p, gamma, nn * sizeof(double)));
    hipLaunchKernelGGL(la_norm_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream, dG.as<double>(), n, dN.as<double>());
    hipLaunchKernelGGL(la_scale_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, dG.as<double>(), n,
                       dN.as<double>(), dC.as<double>());
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    int bad = 0;
    MCMB_TRY(la_potrf(c, dC.as<double>(), n, dF.as<int>(), &bad));
    if (bad_pivot) *bad_pivot = bad ? bad - 1 : -1;
    if (bad) return fail(MCMB_ERR_NOT_SPD, "Cholesky: non-positive pivot at column %d of %d (normalised Gram not "
                                           "positive definite in FP64)", bad - 1, n);
    // Linv = L^-1: solve L X = I column by column (RHS-major X)
    MCMB_TRY(dX.alloc(nn * sizeof(double)));
    hipLaunchKernelGGL(la_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, dX.as<double>(), n);
    MCMB_TRY(la_solve(c, dC.as<double>(), n, 0, dX.as<double>(), n, true));
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_TRY(dO.alloc(nn * sizeof(double)));
    MCMB_TRY(dL.alloc(nn * sizeof(double)));
    hipLaunchKernelGGL(la_outputs, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, dC.as<double>(),
                       dX.as<double>(), dN.as<double>(), n, dO.as<double>(), dL.as<double>());
    MCMB_HIP(hipGetLastError());
    if (ortho) MCMB_TRY(copy_to_host(c, ortho, dO.p, nn * sizeof(double)));
    if (lambda) MCMB_TRY(copy_to_host(c, lambda, dL.p, nn * sizeof(double)));
    if (norm) MCMB_TRY(copy_to_host(c, norm, dN.p, n * sizeof(double)));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 0;
    return MCMB_OK;
}

extern "C" int mcmb_cholesky_solve(mcmb_ctx* c, int n, const double* ortho, int nrhs, const double* alpha,
                                   double* beta)
{
    if (!c || !ortho || !alpha || !beta || n < 1 || nrhs < 1) return fail(MCMB_ERR_ARG, "mcmb_cholesky_solve: bad argument");
    MCMB_TRY(set_device(c));
    const size_t nn = (size_t)n * n;
    DevBuf dF, dX;
    MCMB_TRY(dF.alloc(nn * sizeof(double)));
    MCMB_TRY(dX.alloc((size_t)n * nrhs * sizeof(double)));
    MCMB_TRY(copy_to_device(c, dF.p, ortho, nn * sizeof(double)));
    MCMB_TRY(copy_to_device(c, dX.p, alpha, (size_t)n * nrhs * sizeof(double)));
    MCMB_TRY(la_solve(c, dF.as<double>(), n, 0, dX.as<double>(), nrhs));
    MCMB_TRY(la_solve(c, dF.as<double>(), n, 1, dX.as<double>(), nrhs));
    MCMB_TRY(copy_to_host(c, beta, dX.p, (size_t)n * nrhs * sizeof(double)));
    return MCMB_OK;
}
