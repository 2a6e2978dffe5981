// FP64 orthonormalisation on the GPU (ortho_prim_mpi.c:198-284) and the
// Cholesky solve of the projection (eigenmode_mpi.c:195-219).
//
//   norm_i = sqrt(G_ii),  C = D^-1 G D^-1 = L L^T  (blocked right-looking
//   Cholesky, block NB: one-workgroup LDS factorisation of the diagonal block,
//   row-parallel triangular solve of the panel, tiled SYRK update of the
//   trailing lower triangle), Linv = L^-1 by blocked forward substitution.
//   ortho  = symmetrised D L          (update_ortho, fixed_data.c:3285-3295)
//   lambda = symmetrised (D L)^-1     (update_lambda, fixed_data.c:3576-3586)
// A non-positive pivot sets a device flag (first failing column + 1); later
// kernels see it and return, and the host reports MCMB_ERR_NOT_SPD.
#include "mcmb_internal.h"
#include <vector>

namespace mcmb {

constexpr int NB = 32;     // Cholesky / substitution block
constexpr int GT = 64;     // SYRK / GEMM output tile

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

// Trailing update C[i][j] -= sum_l L[i][k0+l] L[j][k0+l] on lower tiles of
// [k1, n)^2 (k1 = k0 + nb); tile list = (I, J) with I >= J, GT x GT.
__global__ __launch_bounds__(256) void la_syrk(double* __restrict__ C, int n, int k0, int k1,
                                               const int* __restrict__ flag)
{
    __shared__ double Pi[GT][NB + 1], Pj[GT][NB + 1];
    if (*flag) return;
    const int nb = k1 - k0;
    // decode lower tile index t -> (I, J), I >= J
    const int t = blockIdx.x;
    int I = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((I + 1) * (I + 2) / 2 <= t) ++I;
    while (I * (I + 1) / 2 > t) --I;
    const int J = t - I * (I + 1) / 2;
    const int i0 = k1 + I * GT, j0 = k1 + J * GT;
    for (int e = threadIdx.x; e < GT * NB; e += 256) {
        int r = e / NB, l = e % NB;
        Pi[r][l] = (i0 + r < n && l < nb) ? C[(size_t)(i0 + r) * n + k0 + l] : 0.0;
        Pj[r][l] = (j0 + r < n && l < nb) ? C[(size_t)(j0 + r) * n + k0 + l] : 0.0;
    }
    __syncthreads();
    const int tr = (threadIdx.x >> 4) * 4, tc = (threadIdx.x & 15) * 4;
    double acc[4][4] = {};
    for (int l = 0; l < nb; ++l) {
        double a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) { a[u] = Pi[tr + u][l]; b[u] = Pj[tc + u][l]; }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += a[u] * b[v];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int i = i0 + tr + u, j = j0 + tc + v;
            if (i < n && j < n && j <= i) C[(size_t)i * n + j] -= acc[u][v];
        }
}

// X[r][rb..rb+nb) -= sum_{k < kend} Aop(rb+m, k) X[r][k], RHS-major X (ldx),
// Aop(i,k) = A[i*n + k] (trans == 0) or A[k*n + i] (trans == 1); k range
// [kbeg, kend).  Tile: NB rows x 64 RHS.
__global__ __launch_bounds__(256) void la_gemm_update(const double* __restrict__ A, int n, int trans,
                                                      double* __restrict__ X, int ldx, int nrhs, int rb,
                                                      int kbeg, int kend)
{
    __shared__ double As[NB][NB + 1];
    __shared__ double Bs[64][NB + 1];
    const int r0 = blockIdx.x * 64;
    const int tm = (threadIdx.x >> 5), tr = (threadIdx.x & 31);   // 8 x 32 threads
    double acc[4][2] = {};
    for (int kc = kbeg; kc < kend; kc += NB) {
        for (int e = threadIdx.x; e < NB * NB; e += 256) {
            int m = e / NB, k = e % NB;
            int i = rb + m, kk = kc + k;
            double v = 0.0;
            if (i < n && kk < kend) v = trans ? A[(size_t)kk * n + i] : A[(size_t)i * n + kk];
            As[m][k] = v;
        }
        for (int e = threadIdx.x; e < 64 * NB; e += 256) {
            int r = e / NB, k = e % NB;
            int rr = r0 + r, kk = kc + k;
            Bs[r][k] = (rr < nrhs && kk < kend) ? X[(size_t)rr * ldx + kk] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < NB; ++k) {
            double a[4], b[2];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = As[tm * 4 + u][k];
            b[0] = Bs[tr][k];
            b[1] = Bs[tr + 32][k];
#pragma unroll
            for (int u = 0; u < 4; ++u) { acc[u][0] += a[u] * b[0]; acc[u][1] += a[u] * b[1]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            int i = rb + tm * 4 + u, rr = r0 + tr + 32 * v;
            if (i < n && rr < nrhs && tm * 4 + u < NB) X[(size_t)rr * ldx + i] -= acc[u][v];
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

// Forward (L y = b) or backward (L^T y = b) substitution on nrhs RHS-major columns.
static int la_solve(mcmb_ctx* c, const double* dL, int n, int upper, double* dX, int nrhs)
{
    const unsigned gr = (unsigned)((nrhs + 63) / 64), gt = (unsigned)((nrhs + 255) / 256);
    const int nblk = (n + NB - 1) / NB;
    for (int b = 0; b < nblk; ++b) {
        const int kb = upper ? (nblk - 1 - b) : b;
        const int rb = kb * NB;
        if (!upper && rb > 0)
            hipLaunchKernelGGL(la_gemm_update, dim3(gr), dim3(256), 0, c->stream, dL, n, 0, dX, n, nrhs, rb, 0, rb);
        if (upper && rb + NB < n)
            hipLaunchKernelGGL(la_gemm_update, dim3(gr), dim3(256), 0, c->stream, dL, n, 1, dX, n, nrhs, rb, rb + NB, n);
        hipLaunchKernelGGL(la_trsm_rhs, dim3(gt), dim3(256), 0, c->stream, dL, n, upper, dX, n, nrhs, rb);
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
            const int T = (n - k1 + GT - 1) / GT;
            hipLaunchKernelGGL(la_syrk, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, c->stream, dC, n, k0, k1,
                               d_flag);
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
    MCMB_HIP(hipMemcpyAsync(dG.p, gamma, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
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
    MCMB_TRY(la_solve(c, dC.as<double>(), n, 0, dX.as<double>(), n));
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_TRY(dO.alloc(nn * sizeof(double)));
    MCMB_TRY(dL.alloc(nn * sizeof(double)));
    hipLaunchKernelGGL(la_outputs, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, dC.as<double>(),
                       dX.as<double>(), dN.as<double>(), n, dO.as<double>(), dL.as<double>());
    MCMB_HIP(hipGetLastError());
    if (ortho) MCMB_HIP(hipMemcpyAsync(ortho, dO.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (lambda) MCMB_HIP(hipMemcpyAsync(lambda, dL.p, nn * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (norm) MCMB_HIP(hipMemcpyAsync(norm, dN.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
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
    MCMB_HIP(hipMemcpyAsync(dF.p, ortho, nn * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(dX.p, alpha, (size_t)n * nrhs * sizeof(double), hipMemcpyHostToDevice, c->stream));
    MCMB_TRY(la_solve(c, dF.as<double>(), n, 0, dX.as<double>(), nrhs));
    MCMB_TRY(la_solve(c, dF.as<double>(), n, 1, dX.as<double>(), nrhs));
    MCMB_HIP(hipMemcpyAsync(beta, dX.p, (size_t)n * nrhs * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}
