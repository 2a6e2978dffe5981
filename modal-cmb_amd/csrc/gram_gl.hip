// Gauss-Legendre-factorised Gram (calculate_orthol, orthogonalisation_late.c:
// 239-336; M_ij of calculate_orthoMij_T, 99-129), FP64 on the GPU.
//
// With int_{-1}^{1} P_l1 P_l2 P_l3 dx = 2 (l1 l2 l3; 0 0 0)^2, the cube sum of
// h^2 f1(l1) f2(l2) f3(l3) is (1/8 pi) int dx prod_a sum_l (2l+1) P_l(x) f_a(l),
// exact on npt >= (3 lmax + 1)/2 nodes; symmetrising the mode products leaves
// 1/6 of the 6 pairings, so
//   M_ij(x_s) = sum_{l=lmin}^{lmax} (2l+1) v(l) P_l(x_s) q_i(l) q_j(l)
//   Gamma_mn  = sum_s w_s/(48 pi) perm[ M_{m_a n_b}(x_s) ]_{a,b=1..3}
// (perm = permanent of the 3 x 3 matrix).  Three kernels: Legendre table by
// recurrence (one thread per node), M by a tiled FP64 GEMM over l, and the
// node sum per 64 x 64 tile of (m, n) with M staged in LDS 8 nodes at a time.
// O(R^2 npt L + n^2 npt) instead of O(n^2 N_tet).
#include "mcmb_internal.h"
#include <cmath>
#include <mutex>
#include <vector>

namespace mcmb {

constexpr int GLR = 32;   // max basis width
constexpr int SC = 8;     // nodes per LDS stage

// Pt[l][s] = (2l+1) v(l) P_l(x_s) for lmin <= l <= lmax, 0 below lmin.
__global__ __launch_bounds__(256) void gl_legendre(const double* __restrict__ x, int npt, int lmin, int lmax,
                                                   const double* __restrict__ v, double* __restrict__ Pt)
{
    int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= npt) return;
    double xs = x[s], p0 = 1.0, p1 = xs;
    for (int l = 0; l <= lmax; ++l) {
        double pl;
        if (l == 0) pl = 1.0;
        else if (l == 1) pl = xs;
        else {
            pl = ((2.0 * l - 1.0) * xs * p1 - (l - 1.0) * p0) / l;
            p0 = p1;
            p1 = pl;
        }
        Pt[(size_t)l * npt + s] = (l >= lmin) ? (2.0 * l + 1.0) * v[l] * pl : 0.0;
    }
}

// U[l][p] = q_i(l) q_j(l) for the pair p = (i, j), i <= j < R (p enumerates the
// upper triangle row by row); l <= lmax.
__global__ void gl_pairs(const double* __restrict__ q, int ldq, int R, int lmax, double* __restrict__ U, int npair)
{
    int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= (lmax + 1) * npair) return;
    int l = e / npair, p = e % npair;
    int i = 0, rem = p;
    while (rem >= R - i) { rem -= R - i; ++i; }
    int j = i + rem;
    U[e] = q[(size_t)l * ldq + i] * q[(size_t)l * ldq + j];
}

// Mp[p][s] = sum_l U[l][p] Pt[l][s]  (tile 64 s x 32 p, K = l in steps of 16)
__global__ __launch_bounds__(256) void gl_mgemm(const double* __restrict__ Pt, const double* __restrict__ U,
                                                int npt, int npair, int nl, double* __restrict__ Mp)
{
    __shared__ double Ps[16][64 + 1], Us[16][32 + 1];
    const int s0 = blockIdx.x * 64, p0 = blockIdx.y * 32;
    const int ts = (threadIdx.x & 15) * 4, tp = (threadIdx.x >> 4) * 2;
    double acc[4][2] = {};
    for (int l0 = 0; l0 < nl; l0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            int l = e / 64, s = e % 64;
            Ps[l][s] = (l0 + l < nl && s0 + s < npt) ? Pt[(size_t)(l0 + l) * npt + s0 + s] : 0.0;
        }
        for (int e = threadIdx.x; e < 16 * 32; e += 256) {
            int l = e / 32, p = e % 32;
            Us[l][p] = (l0 + l < nl && p0 + p < npair) ? U[(size_t)(l0 + l) * npair + p0 + p] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            double a[4], b[2];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = Ps[l][ts + u];
            b[0] = Us[l][tp];
            b[1] = Us[l][tp + 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) { acc[u][0] += a[u] * b[0]; acc[u][1] += a[u] * b[1]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            int s = s0 + ts + u, p = p0 + tp + v;
            if (s < npt && p < npair) Mp[(size_t)p * npt + s] = acc[u][v];
        }
}

// Gamma over 64 x 64 mode tiles (I <= J): each thread owns a 4 x 4 block of
// (m, n) and sums w_s/(48 pi) perm(M_{m_a n_b}(x_s)) over all nodes; M for SC
// nodes at a time is staged in LDS as Ms[s][i][j] (full R x R).
__global__ __launch_bounds__(256) void gl_gamma(const double* __restrict__ Mp, const double* __restrict__ wgt,
                                                int npt, int R, const int* __restrict__ modes, int n,
                                                const int2* __restrict__ tiles, double* __restrict__ out)
{
    __shared__ double Ms[SC][GLR * GLR];
    __shared__ double ws[SC];
    const int2 IJ = tiles[blockIdx.x];
    const int tm = (threadIdx.x >> 4) * 4, tn = (threadIdx.x & 15) * 4;
    int ma[4][3], nb[4][3];
    bool okm[4], okn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        int m = IJ.x * 64 + tm + u, c = IJ.y * 64 + tn + u;
        okm[u] = m < n;
        okn[u] = c < n;
        int pm = okm[u] ? modes[m] : 0, pn = okn[u] ? modes[c] : 0;
        ma[u][0] = pm & 255; ma[u][1] = (pm >> 8) & 255; ma[u][2] = (pm >> 16) & 255;
        nb[u][0] = pn & 255; nb[u][1] = (pn >> 8) & 255; nb[u][2] = (pn >> 16) & 255;
    }
    double acc[4][4] = {};
    for (int s0 = 0; s0 < npt; s0 += SC) {
        __syncthreads();
        for (int e = threadIdx.x; e < SC * R * R; e += 256) {
            int s = e / (R * R), ij = e % (R * R), i = ij / R, j = ij % R;
            int a = i < j ? i : j, b = i < j ? j : i;
            int p = a * R - a * (a - 1) / 2 + (b - a);
            Ms[s][i * GLR + j] = (s0 + s < npt) ? Mp[(size_t)p * npt + s0 + s] : 0.0;
        }
        if (threadIdx.x < SC) ws[threadIdx.x] = (s0 + threadIdx.x < npt) ? wgt[s0 + threadIdx.x] : 0.0;
        __syncthreads();
        for (int s = 0; s < SC; ++s) {
            const double* M = Ms[s];
            const double w = ws[s];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const double a11 = M[ma[u][0] * GLR + nb[v][0]], a12 = M[ma[u][0] * GLR + nb[v][1]], a13 = M[ma[u][0] * GLR + nb[v][2]];
                    const double a21 = M[ma[u][1] * GLR + nb[v][0]], a22 = M[ma[u][1] * GLR + nb[v][1]], a23 = M[ma[u][1] * GLR + nb[v][2]];
                    const double a31 = M[ma[u][2] * GLR + nb[v][0]], a32 = M[ma[u][2] * GLR + nb[v][1]], a33 = M[ma[u][2] * GLR + nb[v][2]];
                    const double perm = a11 * (a22 * a33 + a23 * a32) + a12 * (a21 * a33 + a23 * a31)
                                      + a13 * (a21 * a32 + a22 * a31);
                    acc[u][v] += w * perm;
                }
        }
    }
    const double scale = 1.0 / (48.0 * M_PI);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            int m = IJ.x * 64 + tm + u, c = IJ.y * 64 + tn + v;
            if (okm[u] && okn[v] && m <= c) {
                out[(size_t)m * n + c] = scale * acc[u][v];
                out[(size_t)c * n + m] = scale * acc[u][v];
            }
        }
}

// Tile list of the node sum: the T (T + 1) / 2 pairs (I <= J) of 64-mode
// tiles, I-major, one thread per tile.
__global__ void gl_tiles(int T, int2* __restrict__ tiles)
{
    int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= T * (T + 1) / 2) return;
    int I = 0;
    while (t >= T - I) { t -= T - I; ++I; }
    tiles[blockIdx.x * 256 + threadIdx.x] = make_int2(I, I + t);
}

// Device buffers of one GL Gram: node table, pair products, M and tiles.
struct GlBufs {
    double *x, *w, *Pt, *U, *M;
    int2* tiles;
};

static size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

// Bytes of the buffers of one GL Gram, each region aligned to 256 bytes.
static size_t gl_bytes(int npt, int nl, int npair, int ntiles)
{
    return 2 * round256(sizeof(double) * (size_t)npt) + round256(sizeof(double) * (size_t)nl * npt) +
           round256(sizeof(double) * (size_t)nl * npair) + round256(sizeof(double) * (size_t)npair * npt) +
           round256(sizeof(int2) * (size_t)ntiles);
}

static GlBufs gl_carve(char* base, int npt, int nl, int npair)
{
    GlBufs b;
    b.x = reinterpret_cast<double*>(base);
    base += round256(sizeof(double) * (size_t)npt);
    b.w = reinterpret_cast<double*>(base);
    base += round256(sizeof(double) * (size_t)npt);
    b.Pt = reinterpret_cast<double*>(base);
    base += round256(sizeof(double) * (size_t)nl * npt);
    b.U = reinterpret_cast<double*>(base);
    base += round256(sizeof(double) * (size_t)nl * npair);
    b.M = reinterpret_cast<double*>(base);
    base += round256(sizeof(double) * (size_t)npair * npt);
    b.tiles = reinterpret_cast<int2*>(base);
    return b;
}

// The node table x, w (host, npt entries) is copied in, the four kernels run
// on the context stream, and the call waits for them (x and w may live on the
// caller's stack).
static int gl_run(mcmb_ctx* c, int npt, const double* x, const double* w, const GlBufs& b, double* d_gamma)
{
    LSpace& L = c->ls;
    const int n = L.n, R = L.R, lmax = L.lmax;
    const int npair = R * (R + 1) / 2, nl = lmax + 1;
    const int T = (n + 63) / 64, ntiles = T * (T + 1) / 2;
    MCMB_HIP(hipMemcpyAsync(b.x, x, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(b.w, w, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(gl_tiles, dim3((ntiles + 255) / 256), dim3(256), 0, c->stream, T, b.tiles);
    hipLaunchKernelGGL(gl_legendre, dim3((npt + 255) / 256), dim3(256), 0, c->stream, (const double*)b.x, npt,
                       L.lmin, lmax, L.d_v.as<double>(), b.Pt);
    hipLaunchKernelGGL(gl_pairs, dim3((unsigned)((nl * npair + 255) / 256)), dim3(256), 0, c->stream,
                       L.d_q.as<double>(), L.ldq, R, lmax, b.U, npair);
    hipLaunchKernelGGL(gl_mgemm, dim3((npt + 63) / 64, (npair + 31) / 32), dim3(256), 0, c->stream,
                       (const double*)b.Pt, (const double*)b.U, npt, npair, nl, b.M);
    hipLaunchKernelGGL(gl_gamma, dim3((unsigned)ntiles), dim3(256), 0, c->stream, (const double*)b.M,
                       (const double*)b.w, npt, R, L.d_modes.as<int>(), n, (const int2*)b.tiles, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = 5;
    c->last_rank = 0;   // the direct node sum over the pair products
    return MCMB_OK;
}

int gl_nodes(mcmb_ctx* c, int n, const double** x, const double** w)
{
    if (n < 1) return fail(MCMB_ERR_ARG, "Gauss-Legendre rule of order %d", n);
    if ((int)c->gl_x.size() != n) {
        c->gl_x.assign(n, 0.0);
        c->gl_w.assign(n, 0.0);
        // roots dealt round-robin in blocks of 16 (the cost per root is even)
        const int nr = (n + 1) / 2, nth = std::max(1, std::min(host_threads(), (nr + 15) / 16));
        double* const xs = c->gl_x.data();
        double* const ws = c->gl_w.data();
        run_threads(nth, [&](int t) {
            for (int b = 16 * t; b < nr; b += 16 * nth) mcmb_gl_roots(n, b, std::min(nr, b + 16), xs, ws);
        });
    }
    *x = c->gl_x.data();
    *w = c->gl_w.data();
    return MCMB_OK;
}

static int gl_check(mcmb_ctx* c, int* npt)
{
    LSpace& L = c->ls;
    if (L.R > GLR) return fail(MCMB_ERR_ARG, "GL Gram supports basis width <= %d (got %d)", GLR, L.R);
    if (*npt <= 0) *npt = 3 * L.lmax / 2 + 1;
    return MCMB_OK;
}

int gram_gl_launch(mcmb_ctx* c, int npt, double* d_gamma)
{
    LSpace& L = c->ls;
    MCMB_TRY(gl_check(c, &npt));
    const int n = L.n, R = L.R, nl = L.lmax + 1, npair = R * (R + 1) / 2;
    const int T = (n + 63) / 64;
    const double *x, *w;
    MCMB_TRY(gl_nodes(c, npt, &x, &w));
    // the pair products of Legendre-like bases have a low-rank basis: the
    // node sum then runs over it (r^3 core instead of a permanent per mode
    // pair and node)
    bool used = false;
    MCMB_TRY(gram_gl_compressed(c, npt, x, w, d_gamma, &used));
    if (used) return MCMB_OK;
    DevBuf buf;
    MCMB_TRY(buf.alloc(gl_bytes(npt, nl, npair, T * (T + 1) / 2)));
    return gl_run(c, npt, x, w, gl_carve(buf.as<char>(), npt, nl, npair), d_gamma);
}

// ---- polarised components (calculate_orthol result[1..3]) -----------------

// Mixed components over 64 x 64 mode tiles (I <= J), per node s:
//   KIND 1, TTE, modes (T, T, E): (MT_{a0 b0} MT_{a1 b1} + MT_{a1 b0} MT_{a0 b1}) ME_{a2 b2}
//   KIND 2, TEE, modes (T, E, E): MT_{a0 b0} (ME_{a1 b1} ME_{a2 b2} + ME_{a2 b1} ME_{a1 b2})
// summed with w_s / (16 pi) (orthogonalisation_late.c:302-316).  As
// gl_gamma, with both fields' M staged in LDS GS nodes at a time.
constexpr int GS = 3;
template <int KIND>
__global__ __launch_bounds__(256) void gl_gamma_mixed(const double* __restrict__ MT, int RT,
                                                      const double* __restrict__ ME, int RE,
                                                      const double* __restrict__ wgt, int npt,
                                                      const int* __restrict__ modes, int n,
                                                      const int2* __restrict__ tiles, double* __restrict__ out)
{
    __shared__ double Ts[GS][GLR * GLR], Es[GS][GLR * GLR];
    __shared__ double ws[GS];
    const int2 IJ = tiles[blockIdx.x];
    const int tm = (threadIdx.x >> 4) * 4, tn = (threadIdx.x & 15) * 4;
    int ma[4][3], nb[4][3];
    bool okm[4], okn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int m = IJ.x * 64 + tm + u, c = IJ.y * 64 + tn + u;
        okm[u] = m < n;
        okn[u] = c < n;
        const int pm = okm[u] ? modes[m] : 0, pn = okn[u] ? modes[c] : 0;
        ma[u][0] = pm & 255; ma[u][1] = (pm >> 8) & 255; ma[u][2] = (pm >> 16) & 255;
        nb[u][0] = pn & 255; nb[u][1] = (pn >> 8) & 255; nb[u][2] = (pn >> 16) & 255;
    }
    // stage R x R of the pair-packed Mp (pairs of one field enumerated as in gl_pairs)
    auto stage = [&](double (*S)[GLR * GLR], const double* Mp, int R, int s0) {
        for (int e = threadIdx.x; e < GS * R * R; e += 256) {
            const int s = e / (R * R), ij = e % (R * R), i = ij / R, j = ij % R;
            const int a = i < j ? i : j, b = i < j ? j : i;
            const int p = a * R - a * (a - 1) / 2 + (b - a);
            S[s][i * GLR + j] = (s0 + s < npt) ? Mp[(size_t)p * npt + s0 + s] : 0.0;
        }
    };
    double acc[4][4] = {};
    for (int s0 = 0; s0 < npt; s0 += GS) {
        __syncthreads();
        stage(Ts, MT, RT, s0);
        stage(Es, ME, RE, s0);
        if (threadIdx.x < GS) ws[threadIdx.x] = (s0 + threadIdx.x < npt) ? wgt[s0 + threadIdx.x] : 0.0;
        __syncthreads();
        for (int s = 0; s < GS; ++s) {
            const double* A = Ts[s];
            const double* B = Es[s];
            const double w = ws[s];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    double t;
                    if (KIND == 1) {
                        t = (A[ma[u][0] * GLR + nb[v][0]] * A[ma[u][1] * GLR + nb[v][1]] +
                             A[ma[u][1] * GLR + nb[v][0]] * A[ma[u][0] * GLR + nb[v][1]]) *
                            B[ma[u][2] * GLR + nb[v][2]];
                    } else {
                        t = A[ma[u][0] * GLR + nb[v][0]] *
                            (B[ma[u][1] * GLR + nb[v][1]] * B[ma[u][2] * GLR + nb[v][2]] +
                             B[ma[u][2] * GLR + nb[v][1]] * B[ma[u][1] * GLR + nb[v][2]]);
                    }
                    acc[u][v] += w * t;
                }
        }
    }
    const double scale = 1.0 / (16.0 * M_PI);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = IJ.x * 64 + tm + u, c = IJ.y * 64 + tn + v;
            if (okm[u] && okn[v] && m <= c) {
                out[(size_t)m * n + c] = scale * acc[u][v];
                out[(size_t)c * n + m] = scale * acc[u][v];
            }
        }
}

// One field of the polarised Gram: Legendre table, pair products and M over
// [lmin, lmax] with basis q[l][ldq] (first R columns) and per-leg weight v.
struct GlField {
    int lmin, lmax, ldq, R;
    const double* q;
    const double* v;
};

// Node count of the polarised Gram: 3 max(T lmax, E lmax) / 2 + 1, as
// init_orthol_glint takes it with polarisation on (orthogonalisation_late.c:33-36).
int gl_pol_npt(const mcmb_ctx* c)
{
    int lmax = c->ls.lmax;
    if (c->es.ready && c->es.lmax > lmax) lmax = c->es.lmax;
    return 3 * lmax / 2 + 1;
}

// Components TTE (1), TEE (2), EEE (3) into the device buffer d_gamma (n x n,
// n = the E setup's mode count).
int gram_gl_pol_launch(mcmb_ctx* c, int pol, int npt, double* d_gamma)
{
    const PolSet& E = c->es;
    const LSpace& L = c->ls;
    const std::vector<int>& ord = E.order[pol - 1];
    // slot field per component (0 = T, 1 = E) and the width each field needs
    static const int field[3][3] = {{0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    int R[2] = {0, 0};
    for (int m = 0; m < E.n; ++m)
        for (int a = 0; a < 3; ++a) {
            const int f = field[pol - 1][a], p = ord[3 * m + a];
            if (f == 0 && p >= L.ldq)
                return fail(MCMB_ERR_ORDER, "polarised Gram: mode %d uses T basis column %d of %d", m, p, L.ldq);
            R[f] = std::max(R[f], p + 1);
        }
    const GlField F[2] = {{L.lmin, L.lmax, L.ldq, R[0], L.d_q.as<double>(), L.d_v.as<double>()},
                          {E.lmin, E.lmax, E.ldq, R[1], E.d_q.as<double>(), E.d_v.as<double>()}};
    const int n = E.n, T = (n + 63) / 64, ntiles = T * (T + 1) / 2;
    const double *x, *w;
    MCMB_TRY(gl_nodes(c, npt, &x, &w));
    {
        // over the compressed bases when their pair products have a low rank
        bool used = false;
        if (pol == 3) MCMB_TRY(gram_gl_eee_compressed(c, R[1], npt, x, w, d_gamma, &used));
        else MCMB_TRY(gram_gl_mixed_compressed(c, pol, R[0], R[1], npt, x, w, d_gamma, &used));
        if (used) return MCMB_OK;
    }
    c->last_rank = 0;
    // x, w, then per used field Pt, U, M; then the tile list
    size_t bytes = 2 * round256(sizeof(double) * npt) + round256(sizeof(int2) * (size_t)ntiles);
    for (int f = 0; f < 2; ++f)
        if (F[f].R > 0) {
            const size_t nl = F[f].lmax + 1, np = (size_t)F[f].R * (F[f].R + 1) / 2;
            bytes += round256(sizeof(double) * nl * npt) + round256(sizeof(double) * nl * np) +
                     round256(sizeof(double) * np * npt);
        }
    DevBuf buf;
    MCMB_TRY(buf.alloc(bytes));
    char* p = buf.as<char>();
    double* dx = reinterpret_cast<double*>(p);
    p += round256(sizeof(double) * npt);
    double* dw = reinterpret_cast<double*>(p);
    p += round256(sizeof(double) * npt);
    int2* dtiles = reinterpret_cast<int2*>(p);
    p += round256(sizeof(int2) * (size_t)ntiles);
    MCMB_HIP(hipMemcpyAsync(dx, x, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(dw, w, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(gl_tiles, dim3((ntiles + 255) / 256), dim3(256), 0, c->stream, T, dtiles);
    const double* Mf[2] = {nullptr, nullptr};
    for (int f = 0; f < 2; ++f) {
        if (F[f].R == 0) continue;
        const int nl = F[f].lmax + 1, np = F[f].R * (F[f].R + 1) / 2;
        double* Pt = reinterpret_cast<double*>(p);
        p += round256(sizeof(double) * (size_t)nl * npt);
        double* U = reinterpret_cast<double*>(p);
        p += round256(sizeof(double) * (size_t)nl * np);
        double* M = reinterpret_cast<double*>(p);
        p += round256(sizeof(double) * (size_t)np * npt);
        hipLaunchKernelGGL(gl_legendre, dim3((npt + 255) / 256), dim3(256), 0, c->stream, (const double*)dx, npt,
                           F[f].lmin, F[f].lmax, F[f].v, Pt);
        hipLaunchKernelGGL(gl_pairs, dim3((unsigned)((nl * np + 255) / 256)), dim3(256), 0, c->stream, F[f].q,
                           F[f].ldq, F[f].R, F[f].lmax, U, np);
        hipLaunchKernelGGL(gl_mgemm, dim3((npt + 63) / 64, (np + 31) / 32), dim3(256), 0, c->stream,
                           (const double*)Pt, (const double*)U, npt, np, nl, M);
        Mf[f] = M;
    }
    const int* modes = E.d_modes[pol - 1].as<int>();
    if (pol == 3)
        hipLaunchKernelGGL(gl_gamma, dim3((unsigned)ntiles), dim3(256), 0, c->stream, Mf[1], (const double*)dw, npt,
                           R[1], modes, n, (const int2*)dtiles, d_gamma);
    else if (pol == 1)
        hipLaunchKernelGGL(gl_gamma_mixed<1>, dim3((unsigned)ntiles), dim3(256), 0, c->stream, Mf[0], R[0], Mf[1],
                           R[1], (const double*)dw, npt, modes, n, (const int2*)dtiles, d_gamma);
    else
        hipLaunchKernelGGL(gl_gamma_mixed<2>, dim3((unsigned)ntiles), dim3(256), 0, c->stream, Mf[0], R[0], Mf[1],
                           R[1], (const double*)dw, npt, modes, n, (const int2*)dtiles, d_gamma);
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return MCMB_OK;
}

// ---- the allocation-free variant ----------------------------------------
// Scratch of mcmb_gamma_gl_noalloc: one array per GPU, held until the process
// ends, shared by every call on that GPU (calls that overlap on one GPU race).
// 128 MiB covers the GL Gram up to lmax 2500 with basis width 32 (3751 nodes).
constexpr size_t GL_SCRATCH_BYTES = (size_t)128 << 20;
// Node count whose node table fits the host stack of the call.
constexpr int GL_NPT_MAX = 4096;
static __device__ __align__(256) double scratch[GL_SCRATCH_BYTES / sizeof(double)];

constexpr int GL_MAX_DEVICES = 64;
static void* g_scratch[GL_MAX_DEVICES] = {};

// Device address of this GPU's scratch array, looked up once per device with
// that device current (the array's name is a host-side stand-in, not a
// device address).  Called from mcmb_create.
int gl_scratch_setup(int device)
{
    if (device < 0 || device >= GL_MAX_DEVICES)
        return fail(MCMB_ERR_ARG, "device %d beyond the %d the scratch table holds", device, GL_MAX_DEVICES);
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (g_scratch[device]) return MCMB_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        return fail(MCMB_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = ptr;
    return MCMB_OK;
}

}  // namespace mcmb

using namespace mcmb;

extern "C" int mcmb_gamma_gl(mcmb_ctx* c, int npt, double* gamma_host, double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_gl: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_gl: call mcmb_lspace_set first");
    if (npt > 0 && 2 * npt - 1 < 3 * c->ls.lmax)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_gl: npt %d < (3 lmax + 1)/2 is not exact", npt);
    MCMB_TRY(set_device(c));
    const int n = c->ls.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_TRY(gram_gl_launch(c, npt, dst));
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

extern "C" int mcmb_gamma_gl_pol(mcmb_ctx* c, int pol, int npt, double* gamma_host, double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_pol: null context");
    if (pol < MCMB_POL_TTT || pol > MCMB_POL_EEE) return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_pol: pol %d", pol);
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_gl_pol: call mcmb_lspace_set first");
    if (pol != MCMB_POL_TTT && !c->es.ready)
        return fail(MCMB_ERR_STATE, "mcmb_gamma_gl_pol: call mcmb_lspace_set_E first");
    // degree of the node integrand: the sum of the three legs' lmax
    const int lt = c->ls.lmax, le = c->es.ready ? c->es.lmax : lt;
    const int deg = pol == MCMB_POL_TTT ? 3 * lt : pol == MCMB_POL_TTE ? 2 * lt + le
                  : pol == MCMB_POL_TEE ? lt + 2 * le : 3 * le;
    if (npt <= 0) npt = gl_pol_npt(c);
    if (2 * npt - 1 < deg)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_pol: %d nodes are not exact for degree %d", npt, deg);
    MCMB_TRY(set_device(c));
    const int n = pol == MCMB_POL_TTT ? c->ls.n : c->es.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    if (pol == MCMB_POL_TTT) MCMB_TRY(gram_gl_launch(c, npt, dst));
    else MCMB_TRY(gram_gl_pol_launch(c, pol, npt, dst));
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

extern "C" int mcmb_gamma_gl_noalloc(mcmb_ctx* c, int npt, double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_noalloc: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_gl_noalloc: call mcmb_lspace_set first");
    if (!gamma_dev) return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_noalloc: gamma_dev (n x n device doubles) is required");
    if (npt > 0 && 2 * npt - 1 < 3 * c->ls.lmax)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_noalloc: npt %d < (3 lmax + 1)/2 is not exact", npt);
    MCMB_TRY(set_device(c));
    MCMB_TRY(gl_check(c, &npt));
    if (npt > GL_NPT_MAX)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_noalloc: %d nodes exceed the limit of %d", npt, GL_NPT_MAX);
    const LSpace& L = c->ls;
    const int n = L.n, R = L.R, nl = L.lmax + 1, npair = R * (R + 1) / 2;
    const int T = (n + 63) / 64;
    const size_t need = gl_bytes(npt, nl, npair, T * (T + 1) / 2);
    if (need > GL_SCRATCH_BYTES)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_gl_noalloc: lmax %d, basis width %d, %d nodes, %d modes need %zu bytes "
                    "of scratch, above the limit of %zu", L.lmax, R, npt, n, need, GL_SCRATCH_BYTES);
    char* base = static_cast<char*>(c->device >= 0 && c->device < GL_MAX_DEVICES ? g_scratch[c->device] : nullptr);
    if (!base) return fail(MCMB_ERR_STATE, "mcmb_gamma_gl_noalloc: no scratch on device %d", c->device);
    double x[GL_NPT_MAX], w[GL_NPT_MAX];
    MCMB_TRY(mcmb_gauss_legendre(npt, x, w));
    return gl_run(c, npt, x, w, gl_carve(base, npt, nl, npair), gamma_dev);
}
