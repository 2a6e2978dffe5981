// Cross-Gamma of the late-time basis against the projected primordial basis
// q-tilde, FP64 on the GPU, in the two forms the reference computes it:
//
// * 3D (gamma_3d_host -> calculate_gamma_TTT_3D_nested, calculate_gamma_3d.c:
//   274-440 with calculate_sijk_TTT 813-860):
//     Gamma_mn = 6 dphi^2 sum_{l in T(2, Lq)} z(l) Q_m(l) X_n(l),
//     z = permsix h^2 lw(l1) lw(l2) lw(l3),  lw(l) = (2l+1)^(-1/3),
//     X_n(l) = trapezium over x of y_x, y_x = x^2 Sym_n(q3(l, x)),
//     q3[l][p][x] = (2l+1)^(1/6) / sqrt(C_l + N_l / b_l^2) qt[p][l][x],
//   with y_x set to 0 for x < 10 when |y_x| > 1e-18, Q_m the late mode
//   products (0 unless every leg is in the late range).
// * Gauss-Legendre (calculate_gamma_2d / calculate_gamma, calculate_gamma.c:
//   48-303, 511-555, 772-1016):
//     M[x][p][q][s] = sum_l P_l(mu_s) TTw(l) qt[p][l][x] b_q(l),
//     TTw(l) = (2l+1)^(5/6) b_l / sqrt(b_l^2 C_l + N_l),
//     y_x = x^2 sum_s w_s perm3(M[x][p_a][q_b][s]) with the same rule for
//     x < 10, Gamma_mn = dphi^2 / (8 pi) * spline integral of y over x.
//
// The x integrals are linear functionals of the masked samples y_x: the
// trapezium rule and the two cubic-spline rules (end conditions of Forsythe,
// Malcolm and Moler in integrate_cubic_spline, 845-934; natural ends of
// gsl_interp_cspline) are applied as per-sample weights c_x computed once on
// the host, and the rule on x < 10 is applied per (m, n, x) before weighting.
//
// 3D form on the GPU: the walk of T(2, Lq) is cut into chunks of up to 16
// points of one row (fixed l1, l2); per chunk, one kernel forms X_n for all
// primordial modes (the q3 rows of l1, l2 and the 16 l3 values staged in LDS
// per x-chunk; the mode's pair sums of the l1, l2 rows shared by the four
// points of a thread) and another z Q_m for the late modes; a split-K FP64
// MFMA GEMM contracts them over the points.  GL form: M by a tiled GEMM over
// l per (p, q), then per 64 x 64 tile of (m, n) the node sum with M staged in
// LDS 8 nodes at a time, x outermost.
#include "mcmb_internal.h"
#include "weights.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace mcmb {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int XR = 32;       // max primordial / late basis width
constexpr int XC = 8;        // x samples per LDS stage (3D B side)
constexpr int CH = 16;       // points per chunk (3D form)
constexpr int GS = 8;        // nodes per LDS stage (GL form)

// ---- per-sample weights of the x rules (host) ------------------------------

// Solve the tridiagonal system (sub a, diag d, super c) in place with partial
// pivoting (one extra super-diagonal of fill); rhs b becomes the solution.
static bool tridiag_solve(std::vector<double> a, std::vector<double> d, std::vector<double> c, std::vector<double>& b)
{
    const int n = (int)d.size();
    std::vector<double> e(n, 0.0);   // second super-diagonal from row swaps
    for (int i = 0; i < n - 1; ++i) {
        if (std::fabs(a[i]) > std::fabs(d[i])) {   // swap rows i and i+1
            std::swap(d[i], a[i]);
            std::swap(c[i], d[i + 1]);
            if (i + 1 < n - 1) std::swap(e[i], c[i + 1]);
            std::swap(b[i], b[i + 1]);
        }
        if (d[i] == 0.0) return false;
        const double f = a[i] / d[i];
        d[i + 1] -= f * c[i];
        if (i + 1 < n - 1) c[i + 1] -= f * e[i];
        b[i + 1] -= f * b[i];
    }
    if (d[n - 1] == 0.0) return false;
    b[n - 1] /= d[n - 1];
    if (n > 1) b[n - 2] = (b[n - 2] - c[n - 2] * b[n - 1]) / d[n - 2];
    for (int i = n - 3; i >= 0; --i) b[i] = (b[i] - c[i] * b[i + 1] - e[i] * b[i + 2]) / d[i];
    return true;
}

// c_x with sum_x c_x y_x = integral over [x_0, x_{n-1}] of the cubic spline
// through (x, y): the spline integral is t.y - g.M with t_j = (h_{j-1} + h_j)/2,
// g_j = (h_{j-1}^3 + h_j^3)/24 and knot second derivatives M = A^-1 B y, so
// c = t - B^T z with A^T z = g.  Interior rows of A M = B y are the
// continuity equations h_{i-1} M_{i-1} + 2 (h_{i-1} + h_i) M_i + h_i M_{i+1} =
// 6 (dy_i/h_i - dy_{i-1}/h_{i-1}); end rows: M = 0 (natural) or the third
// derivative on the end interval, (M_1 - M_0)/h_0, equal to 6 times the third
// divided difference of the first four points (and the last four).
int spline_weights(const std::vector<double>& x, int kind, std::vector<double>& cw)
{
    const int n = (int)x.size();
    cw.assign(n, 0.0);
    if (kind == 2) {   // trapezium
        for (int i = 0; i + 1 < n; ++i) {
            const double h = x[i + 1] - x[i];
            cw[i] += 0.5 * h;
            cw[i + 1] += 0.5 * h;
        }
        return MCMB_OK;
    }
    if (n < (kind == MCMB_SPLINE_FMM ? 4 : 2))
        return fail(MCMB_ERR_ARG, "spline rule %d needs at least %d samples (got %d)", kind,
                    kind == MCMB_SPLINE_FMM ? 4 : 2, n);
    std::vector<double> h(n - 1);
    for (int i = 0; i + 1 < n; ++i) {
        h[i] = x[i + 1] - x[i];
        if (!(h[i] > 0.0)) return fail(MCMB_ERR_ARG, "x samples must increase (x[%d] = %g, x[%d] = %g)", i, x[i],
                                       i + 1, x[i + 1]);
    }
    // A (tridiagonal) and B (rows of up to 4 entries) of A M = B y
    std::vector<double> sub(n - 1, 0.0), dia(n, 0.0), sup(n - 1, 0.0);
    std::vector<std::vector<std::pair<int, double>>> B(n);
    for (int i = 1; i + 1 < n; ++i) {
        sub[i - 1] = h[i - 1];
        dia[i] = 2.0 * (h[i - 1] + h[i]);
        sup[i] = h[i];
        B[i] = {{i - 1, 6.0 / h[i - 1]}, {i, -6.0 / h[i - 1] - 6.0 / h[i]}, {i + 1, 6.0 / h[i]}};
    }
    auto dd3 = [&](int j0) {   // coefficients of the third divided difference over x[j0 .. j0+3]
        std::vector<std::pair<int, double>> r;
        for (int j = j0; j < j0 + 4; ++j) {
            double den = 1.0;
            for (int k = j0; k < j0 + 4; ++k)
                if (k != j) den *= x[j] - x[k];
            r.push_back({j, 6.0 / den});
        }
        return r;
    };
    if (kind == MCMB_SPLINE_NATURAL) {
        dia[0] = 1.0;
        dia[n - 1] = 1.0;
    } else {
        dia[0] = -1.0 / h[0];
        sup[0] = 1.0 / h[0];
        B[0] = dd3(0);
        sub[n - 2] = -1.0 / h[n - 2];
        dia[n - 1] = 1.0 / h[n - 2];
        B[n - 1] = dd3(n - 4);
    }
    if (n == 2 && kind == MCMB_SPLINE_NATURAL) sup[0] = sub[0] = 0.0;
    // A^T z = g: A^T has sub-diagonal sup and super-diagonal sub
    std::vector<double> g(n, 0.0);
    for (int j = 0; j < n; ++j) {
        const double a = j > 0 ? h[j - 1] : 0.0, b = j + 1 < n ? h[j] : 0.0;
        g[j] = (a * a * a + b * b * b) / 24.0;
    }
    if (!tridiag_solve(sup, dia, sub, g)) return fail(MCMB_ERR_ARG, "spline system of %d samples is singular", n);
    for (int j = 0; j < n; ++j) {
        const double a = j > 0 ? h[j - 1] : 0.0, b = j + 1 < n ? h[j] : 0.0;
        cw[j] = 0.5 * (a + b);
    }
    for (int i = 0; i < n; ++i)
        for (const auto& e : B[i]) cw[e.first] -= g[i] * e.second;
    return MCMB_OK;
}

// ---- 3D form ---------------------------------------------------------------

struct Chunk {
    int i, j, k0, cnt;   // points (i, j, k0 + 2 t), t < cnt
};

// q3[l][x][p] = s4(l) qt[p][l][x] (primordial index fastest: the B-side kernel
// stages rows of one l over an x range)
__global__ void cx_q3(const double* __restrict__ qt, const double* __restrict__ s4, int np, int nl, int nx,
                      double* __restrict__ q3)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long)nl * nx * np) return;
    const int p = (int)(e % np);
    const long lx = e / np;
    const int x = (int)(lx % nx), l = (int)(lx / nx);
    q3[e] = s4[l] * qt[((long)p * nl + l) * nx + x];
}

// B[c * 16 + t][n] = sum_x c_x y_x for point t of chunk c and primordial mode
// n (0 for t >= cnt): y_x = x^2 (1/6) [S_rs q3(l3, t) + S_rt q3(l3, s) +
// S_st q3(l3, r)] with S_ab = q3(l1, a) q3(l2, b) + q3(l1, b) q3(l2, a),
// zeroed for x < 10 when |y_x| > 1e-18.  One workgroup per chunk; per tile
// of 64 modes, lane = mode, wave w = points 4w .. 4w+3.
__global__ __launch_bounds__(256) void cx3_bside(const Chunk* __restrict__ chunks, const double* __restrict__ q3,
                                                 int np, int nx, const double* __restrict__ xv,
                                                 const double* __restrict__ cx, const int* __restrict__ modes, int n,
                                                 double* __restrict__ B, int ldb)
{
    __shared__ double qs[2 + CH][XC][XR];
    const Chunk ch = chunks[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long row0 = (long)blockIdx.x * CH;
    const long ldrow = (long)nx * np;
    for (int m0 = 0; m0 < n; m0 += 64) {
        const int m = m0 + lane;
        int r = 0, s = 0, t = 0;
        if (m < n) {
            const int pm = modes[m];
            r = pm & 255;
            s = (pm >> 8) & 255;
            t = (pm >> 16) & 255;
        }
        double acc[4] = {0.0, 0.0, 0.0, 0.0};
        for (int x0 = 0; x0 < nx; x0 += XC) {
            const int xn = min(XC, nx - x0);
            __syncthreads();
            // stage rows l1, l2 and the chunk's l3 values over x0 .. x0 + xn
            for (int e = threadIdx.x; e < (2 + CH) * XC * np; e += 256) {
                const int p = e % np, xx = (e / np) % XC, rw = e / (np * XC);
                double v = 0.0;
                if (xx < xn) {
                    int l = -1;
                    if (rw == 0) l = ch.i;
                    else if (rw == 1) l = ch.j;
                    else if (rw - 2 < ch.cnt) l = ch.k0 + 2 * (rw - 2);
                    if (l >= 0) v = q3[(long)l * ldrow + (long)(x0 + xx) * np + p];
                }
                qs[rw][xx][p] = v;
            }
            __syncthreads();
            for (int xx = 0; xx < xn; ++xx) {
                const int xi = x0 + xx;
                const double x2 = xv[xi] * xv[xi], c = cx[xi];
                const double* a = qs[0][xx];
                const double* b = qs[1][xx];
                const double Srs = a[r] * b[s] + a[s] * b[r];
                const double Srt = a[r] * b[t] + a[t] * b[r];
                const double Sst = a[s] * b[t] + a[t] * b[s];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double* k = qs[2 + 4 * w + u][xx];
                    double y = x2 * (Srs * k[t] + Srt * k[s] + Sst * k[r]) * (1.0 / 6.0);
                    if (xi < 10 && fabs(y) > 1e-18) y = 0.0;
                    acc[u] += c * y;
                }
            }
        }
        if (m < n)
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int pt = 4 * w + u;
                B[(row0 + pt) * ldb + m] = pt < ch.cnt ? acc[u] : 0.0;
            }
    }
}

// At[m][c * 16 + t] = z Q_m(l) for late mode m at point t of chunk c (0 for t
// >= cnt), z = permsix h^2 lw(l1) lw(l2) lw(l3), Q_m = (1/6) sum over the 6
// slot permutations of b(l1) b(l2) b(l3).
__global__ __launch_bounds__(256) void cx3_aside(const Chunk* __restrict__ chunks, const double* __restrict__ b,
                                                 int ldq, const int* __restrict__ modes, int n, int asym,
                                                 const double* __restrict__ ftab, const double* __restrict__ lw,
                                                 double* __restrict__ At, long lda)
{
    const Chunk ch = chunks[blockIdx.x];
    const int t = threadIdx.x & 15;
    const int k = ch.k0 + 2 * t;
    const bool ok = t < ch.cnt;
    const double z = ok ? lspace_weight(ch.i, ch.j, k, asym, ftab, lw) : 0.0;
    const double* qi = b + (long)ch.i * ldq;
    const double* qj = b + (long)ch.j * ldq;
    const double* qk = b + (long)(ok ? k : ch.k0) * ldq;
    const long col = (long)blockIdx.x * CH + t;
    for (int m = threadIdx.x >> 4; m < n; m += 16) {
        const int pm = modes[m];
        const int r = pm & 255, s = (pm >> 8) & 255, u = (pm >> 16) & 255;
        const double Q = (qk[u] * (qi[r] * qj[s] + qi[s] * qj[r]) + qk[s] * (qi[r] * qj[u] + qi[u] * qj[r]) +
                          qk[r] * (qi[s] * qj[u] + qi[u] * qj[s])) * (1.0 / 6.0);
        At[(long)m * lda + col] = ok ? z * Q : 0.0;
    }
}

// Split-K FP64 MFMA GEMM: P[split][M][N] = A[M][k range] B[k range][N] over
// the split's share of K (a multiple of 16), 128 x 128 tiles, 4 waves of
// 64 x 64 on v_mfma_f64_16x16x4f64, K chunks of 16 double-buffered in LDS.
__global__ __launch_bounds__(256, 2) void cx_gemm(const double* __restrict__ A, long lda,
                                                  const double* __restrict__ B, long ldb, int M, int N, long K,
                                                  int kper, double* __restrict__ P)
{
    __shared__ double As[2][16][128 + 4];
    __shared__ double Bs[2][16][128 + 4];
    const int tn = (N + 127) / 128;
    const int m0 = (blockIdx.x / tn) * 128, n0 = (blockIdx.x % tn) * 128;
    const long kb = (long)blockIdx.y * kper * 16, ke = min(K, kb + (long)kper * 16);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[a][c] = d4{0.0, 0.0, 0.0, 0.0};
    double ra[8], rb[8];
    auto gload = [&](long k0) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + u * 256;   // 0 .. 2047
            const int ar = e >> 4, ak = e & 15;
            ra[u] = (m0 + ar < M) ? A[(long)(m0 + ar) * lda + k0 + ak] : 0.0;
            const int bk = e >> 7, bc = e & 127;
            rb[u] = (n0 + bc < N) ? B[(k0 + bk) * ldb + n0 + bc] : 0.0;
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int e = tid + u * 256;
            As[buf][e & 15][e >> 4] = ra[u];
            Bs[buf][e >> 7][e & 127] = rb[u];
        }
    };
    if (kb < ke) {
        gload(kb);
        sstore(0);
    }
    __syncthreads();
    const long nk = (ke - kb) / 16;
    for (long kc = 0; kc < nk; ++kc) {
        const int buf = (int)(kc & 1);
        if (kc + 1 < nk) gload(kb + (kc + 1) * 16);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            const int kk = ks * 4 + (lane >> 4);
            double av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) av[a] = As[buf][kk][wm * 64 + a * 16 + (lane & 15)];
#pragma unroll
            for (int c = 0; c < 4; ++c) bv[c] = Bs[buf][kk][wn * 64 + c * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int c = 0; c < 4; ++c) acc[a][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[c], acc[a][c], 0, 0, 0);
        }
        if (kc + 1 < nk) sstore(buf ^ 1);
        __syncthreads();
    }
    double* Pp = P + (long)blockIdx.y * M * N;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int gm = m0 + wm * 64 + a * 16 + 4 * v + (lane >> 4);
                const int gn = n0 + wn * 64 + c * 16 + (lane & 15);
                if (gm < M && gn < N) Pp[(long)gm * N + gn] = acc[a][c][v];
            }
}

// C[e] (+)= scale * sum over splits of P[split][e], splits added in order
__global__ void cx_reduce(const double* __restrict__ P, int nsplit, long MN, double scale, int accumulate,
                          double* __restrict__ C)
{
    const long e = (long)blockIdx.x * 256 + threadIdx.x;
    if (e >= MN) return;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += P[(long)k * MN + e];
    C[e] = accumulate ? C[e] + scale * s : scale * s;
}

// ---- Gauss-Legendre form ---------------------------------------------------

// Pt[l][s] = TTw(l) P_l(mu_s) for lmin <= l <= lmax, 0 below lmin.
__global__ __launch_bounds__(256) void cx_legendre(const double* __restrict__ mu, int npt, int lmin, int lmax,
                                                   const double* __restrict__ tw, double* __restrict__ Pt)
{
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= npt) return;
    const double xs = mu[s];
    double p0 = 1.0, p1 = xs;
    for (int l = 0; l <= lmax; ++l) {
        double pl;
        if (l == 0) pl = 1.0;
        else if (l == 1) pl = xs;
        else {
            pl = ((2.0 * l - 1.0) * xs * p1 - (l - 1.0) * p0) / l;
            p0 = p1;
            p1 = pl;
        }
        Pt[(long)l * npt + s] = l >= lmin ? tw[l] * pl : 0.0;
    }
}

// M[x][p][q][s] = sum_{l = lmin..lmax} qt[p][l][x] b_q(l) Pt[l][s]; one
// workgroup per (64 nodes, 32 samples of x, pair (p, q)), K = l in steps of 16.
__global__ __launch_bounds__(256) void cx_mgemm(const double* __restrict__ qt, int nlq, int nx,
                                                const double* __restrict__ b, int ldq, int Rq,
                                                const double* __restrict__ Pt, int npt, int lmin, int lmax, int np,
                                                double* __restrict__ M)
{
    __shared__ double Ps[16][64 + 1], Qs[16][32 + 1];
    const int s0 = blockIdx.x * 64, x0 = blockIdx.y * 32;
    const int p = blockIdx.z / Rq, q = blockIdx.z % Rq;
    const int ts = (threadIdx.x & 15) * 4, tx = (threadIdx.x >> 4) * 2;
    double acc[4][2] = {};
    const double* qp = qt + (long)p * nlq * nx;
    for (int l0 = lmin; l0 <= lmax; l0 += 16) {
        for (int e = threadIdx.x; e < 16 * 64; e += 256) {
            const int l = e / 64, s = e % 64;
            Ps[l][s] = (l0 + l <= lmax && s0 + s < npt) ? b[(long)(l0 + l) * ldq + q] * Pt[(long)(l0 + l) * npt + s0 + s]
                                                        : 0.0;
        }
        for (int e = threadIdx.x; e < 16 * 32; e += 256) {
            const int l = e / 32, x = e % 32;
            Qs[l][x] = (l0 + l <= lmax && x0 + x < nx) ? qp[(long)(l0 + l) * nx + x0 + x] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            double a[4], c[2];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = Ps[l][ts + u];
            c[0] = Qs[l][tx];
            c[1] = Qs[l][tx + 1];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[u][0] += a[u] * c[0];
                acc[u][1] += a[u] * c[1];
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 2; ++v) {
            const int s = s0 + ts + u, x = x0 + tx + v;
            if (s < npt && x < nx) M[(((long)x * np + p) * Rq + q) * npt + s] = acc[u][v];
        }
}

// Gamma[m][n] for late modes m (rows, q indices) and primordial modes n
// (columns, p indices), one workgroup per 64 x 64 tile, 4 x 4 per thread:
// per x, g = sum_s w_s perm3(M[x][p_a][q_b][s]); y = x^2 g, zeroed for x < 10
// when |y| > 1e-18; Gamma = scale * sum_x c_x y.
__global__ __launch_bounds__(256) void cx_gl_gamma(const double* __restrict__ M, int np, int Rq, int nx, int npt,
                                                   const double* __restrict__ wgt, const double* __restrict__ xv,
                                                   const double* __restrict__ cx, const int* __restrict__ mlate,
                                                   int nl, const int* __restrict__ mprim, int npm, double scale,
                                                   double* __restrict__ out)
{
    __shared__ double Ms[GS][XR * XR];
    __shared__ double ws[GS];
    const int tnp = (npm + 63) / 64;
    const int I = blockIdx.x / tnp, J = blockIdx.x % tnp;
    const int tm = (threadIdx.x >> 4) * 4, tn = (threadIdx.x & 15) * 4;
    int qa[4][3], pa[4][3];
    bool okm[4], okn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int m = I * 64 + tm + u, c = J * 64 + tn + u;
        okm[u] = m < nl;
        okn[u] = c < npm;
        const int qm = okm[u] ? mlate[m] : 0, pn = okn[u] ? mprim[c] : 0;
        qa[u][0] = qm & 255; qa[u][1] = (qm >> 8) & 255; qa[u][2] = (qm >> 16) & 255;
        pa[u][0] = pn & 255; pa[u][1] = (pn >> 8) & 255; pa[u][2] = (pn >> 16) & 255;
    }
    double tot[4][4] = {};
    const int PQ = np * Rq;
    for (int x = 0; x < nx; ++x) {
        double acc[4][4] = {};
        const double* Mx = M + (long)x * PQ * npt;
        for (int s0 = 0; s0 < npt; s0 += GS) {
            __syncthreads();
            for (int e = threadIdx.x; e < GS * PQ; e += 256) {
                const int s = e % GS, pq = e / GS;
                const int p = pq / Rq, q = pq % Rq;
                Ms[s][p * XR + q] = (s0 + s < npt) ? Mx[(long)pq * npt + s0 + s] : 0.0;
            }
            if (threadIdx.x < GS) ws[threadIdx.x] = (s0 + threadIdx.x < npt) ? wgt[s0 + threadIdx.x] : 0.0;
            __syncthreads();
            for (int s = 0; s < GS; ++s) {
                const double* Mt = Ms[s];
                const double w = ws[s];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        // a_ab = M[p_a of n][q_b of m]
                        const double a11 = Mt[pa[v][0] * XR + qa[u][0]], a12 = Mt[pa[v][0] * XR + qa[u][1]],
                                     a13 = Mt[pa[v][0] * XR + qa[u][2]];
                        const double a21 = Mt[pa[v][1] * XR + qa[u][0]], a22 = Mt[pa[v][1] * XR + qa[u][1]],
                                     a23 = Mt[pa[v][1] * XR + qa[u][2]];
                        const double a31 = Mt[pa[v][2] * XR + qa[u][0]], a32 = Mt[pa[v][2] * XR + qa[u][1]],
                                     a33 = Mt[pa[v][2] * XR + qa[u][2]];
                        const double perm = a11 * (a22 * a33 + a23 * a32) + a12 * (a21 * a33 + a23 * a31) +
                                            a13 * (a21 * a32 + a22 * a31);
                        acc[u][v] += w * perm;
                    }
            }
        }
        const double x2 = xv[x] * xv[x], c = cx[x];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double y = x2 * acc[u][v];
                if (x < 10 && fabs(y) > 1e-18) y = 0.0;
                tot[u][v] += c * y;
            }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = I * 64 + tm + u, c = J * 64 + tn + v;
            if (okm[u] && okn[v]) out[(long)m * npm + c] = scale * tot[u][v];
        }
}

// The polarised components TTE (KIND 1) and TEE (KIND 2) of the above: late
// slot b reads M^T or M^E by its field (TTE: T, T, E; TEE: T, E, E), a_ab =
// M^{X_b}[p_a][q_b] (gamma_2d_pt_TTE / _TEE, calculate_gamma.c:557-653);
// both fields' M staged in LDS GS2 nodes at a time.
constexpr int GS2 = 3;
template <int KIND>
__global__ __launch_bounds__(256) void cx_gl_gamma_mixed(const double* __restrict__ MT, int RT,
                                                         const double* __restrict__ ME, int RE, int np, int nx,
                                                         int npt, const double* __restrict__ wgt,
                                                         const double* __restrict__ xv, const double* __restrict__ cx,
                                                         const int* __restrict__ mlate, int nl,
                                                         const int* __restrict__ mprim, int npm, double scale,
                                                         double* __restrict__ out)
{
    __shared__ double Ts[GS2][XR * XR], Es[GS2][XR * XR];
    __shared__ double ws[GS2];
    const int tnp = (npm + 63) / 64;
    const int I = blockIdx.x / tnp, J = blockIdx.x % tnp;
    const int tm = (threadIdx.x >> 4) * 4, tn = (threadIdx.x & 15) * 4;
    int qa[4][3], pa[4][3];
    bool okm[4], okn[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int m = I * 64 + tm + u, c = J * 64 + tn + u;
        okm[u] = m < nl;
        okn[u] = c < npm;
        const int qm = okm[u] ? mlate[m] : 0, pn = okn[u] ? mprim[c] : 0;
        qa[u][0] = qm & 255; qa[u][1] = (qm >> 8) & 255; qa[u][2] = (qm >> 16) & 255;
        pa[u][0] = pn & 255; pa[u][1] = (pn >> 8) & 255; pa[u][2] = (pn >> 16) & 255;
    }
    // LDS table of each late slot
    const double(*S1)[XR * XR] = Ts;
    const double(*S2)[XR * XR] = KIND == 1 ? Ts : Es;
    const double(*S3)[XR * XR] = Es;
    double tot[4][4] = {};
    const int PT = np * RT, PE = np * RE;
    for (int x = 0; x < nx; ++x) {
        double acc[4][4] = {};
        const double* MTx = MT + (long)x * PT * npt;
        const double* MEx = ME + (long)x * PE * npt;
        for (int s0 = 0; s0 < npt; s0 += GS2) {
            __syncthreads();
            for (int e = threadIdx.x; e < GS2 * PT; e += 256) {
                const int s = e % GS2, pq = e / GS2;
                Ts[s][(pq / RT) * XR + pq % RT] = (s0 + s < npt) ? MTx[(long)pq * npt + s0 + s] : 0.0;
            }
            for (int e = threadIdx.x; e < GS2 * PE; e += 256) {
                const int s = e % GS2, pq = e / GS2;
                Es[s][(pq / RE) * XR + pq % RE] = (s0 + s < npt) ? MEx[(long)pq * npt + s0 + s] : 0.0;
            }
            if (threadIdx.x < GS2) ws[threadIdx.x] = (s0 + threadIdx.x < npt) ? wgt[s0 + threadIdx.x] : 0.0;
            __syncthreads();
            for (int s = 0; s < GS2; ++s) {
                const double *A = S1[s], *B = S2[s], *D = S3[s];
                const double w = ws[s];
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const double a11 = A[pa[v][0] * XR + qa[u][0]], a12 = B[pa[v][0] * XR + qa[u][1]],
                                     a13 = D[pa[v][0] * XR + qa[u][2]];
                        const double a21 = A[pa[v][1] * XR + qa[u][0]], a22 = B[pa[v][1] * XR + qa[u][1]],
                                     a23 = D[pa[v][1] * XR + qa[u][2]];
                        const double a31 = A[pa[v][2] * XR + qa[u][0]], a32 = B[pa[v][2] * XR + qa[u][1]],
                                     a33 = D[pa[v][2] * XR + qa[u][2]];
                        const double perm = a11 * (a22 * a33 + a23 * a32) + a12 * (a21 * a33 + a23 * a31) +
                                            a13 * (a21 * a32 + a22 * a31);
                        acc[u][v] += w * perm;
                    }
            }
        }
        const double x2 = xv[x] * xv[x], c = cx[x];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                double y = x2 * acc[u][v];
                if (x < 10 && fabs(y) > 1e-18) y = 0.0;
                tot[u][v] += c * y;
            }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int m = I * 64 + tm + u, c = J * 64 + tn + v;
            if (okm[u] && okn[v]) out[(long)m * npm + c] = scale * tot[u][v];
        }
}

}  // namespace mcmb

using namespace mcmb;

// ---- setup -------------------------------------------------------------------

extern "C" int mcmb_qtilde_set(mcmb_ctx* c, int np, int lmax_q, int nx, const double* qt, const double* xvec,
                               int nmodes, const int* order_prim, const double* cl, const double* nl,
                               const double* bl, double deltaphi)
{
    if (!c || !qt || !xvec || !order_prim || !cl || !nl || !bl) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: null argument");
    if (np < 1 || np > XR) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: np %d not in [1, %d]", np, XR);
    if (lmax_q < 2 || lmax_q > 8190) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: lmax_q %d not in [2, 8190]", lmax_q);
    if (nx < 2) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: nx %d < 2", nx);
    if (nmodes < 1 || nmodes > 65536) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: nmodes %d", nmodes);
    for (int e = 0; e < 3 * nmodes; ++e)
        if (order_prim[e] < 0 || order_prim[e] >= np)
            return fail(MCMB_ERR_ORDER, "mcmb_qtilde_set: order_prim[%d][%d] = %d outside [0, %d)", e / 3, e % 3,
                        order_prim[e], np);
    for (int x = 0; x + 1 < nx; ++x)
        if (!(xvec[x + 1] > xvec[x])) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set: xvec must increase (at %d)", x);
    MCMB_TRY(set_device(c));
    CrossSet& X = c->xs;
    X.ready = false;
    X.pol = false;
    X.d_qc.release();
    X.np = np;
    X.lq = lmax_q;
    X.nx = nx;
    X.n = nmodes;
    X.dphi = deltaphi;
    const size_t nq = (size_t)np * (lmax_q + 1) * nx;
    X.qt.assign(qt, qt + nq);
    X.xvec.assign(xvec, xvec + nx);
    X.order.assign(order_prim, order_prim + 3 * (size_t)nmodes);
    X.cl.assign(cl, cl + lmax_q + 1);
    X.nl.assign(nl, nl + lmax_q + 1);
    X.bl.assign(bl, bl + lmax_q + 1);
    std::vector<int> modes(nmodes);
    for (int m = 0; m < nmodes; ++m)
        modes[m] = order_prim[3 * m] | (order_prim[3 * m + 1] << 8) | (order_prim[3 * m + 2] << 16);
    MCMB_TRY(X.d_qt.alloc(sizeof(double) * nq));
    MCMB_TRY(X.d_x.alloc(sizeof(double) * nx));
    MCMB_TRY(X.d_modes.alloc(sizeof(int) * nmodes));
    MCMB_HIP(hipMemcpyAsync(X.d_qt.p, X.qt.data(), sizeof(double) * nq, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(X.d_x.p, X.xvec.data(), sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(X.d_modes.p, modes.data(), sizeof(int) * nmodes, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    X.generation = next_generation();
    X.ready = true;
    return MCMB_OK;
}

extern "C" int mcmb_qtilde_set_E(mcmb_ctx* c, const double* qt_E, const double* cl_TE, const double* cl_EE,
                                 const double* nl_TE, const double* nl_EE, const double* bl_TE, const double* bl_EE,
                                 double fsky_ratio)
{
    if (!c || !qt_E || !cl_TE || !cl_EE || !nl_TE || !nl_EE || !bl_TE || !bl_EE)
        return fail(MCMB_ERR_ARG, "mcmb_qtilde_set_E: null argument");
    if (!c->xs.ready) return fail(MCMB_ERR_STATE, "mcmb_qtilde_set_E: call mcmb_qtilde_set first");
    if (!(fsky_ratio >= 0.0)) return fail(MCMB_ERR_ARG, "mcmb_qtilde_set_E: fsky ratio %g", fsky_ratio);
    MCMB_TRY(set_device(c));
    CrossSet& X = c->xs;
    X.pol = false;
    const int nlq = X.lq + 1;
    X.cte.assign(cl_TE, cl_TE + nlq);
    X.cee.assign(cl_EE, cl_EE + nlq);
    X.nte.assign(nl_TE, nl_TE + nlq);
    X.nee.assign(nl_EE, nl_EE + nlq);
    X.bte.assign(bl_TE, bl_TE + nlq);
    X.bee.assign(bl_EE, bl_EE + nlq);
    X.fsky_ratio = fsky_ratio;
    // qc[p][l][x] = s1 qt_E - s2 qt_T, s1 = b_TT^2 C_TT + N_TT,
    // s2 = sqrt(fsky_E / fsky_T) b_TE^2 C_TE + N_TE (init_gamma_glint,
    // calculate_gamma.c:171-183)
    const size_t nq = (size_t)X.np * nlq * X.nx;
    std::vector<double> qc(nq);
    const double rf = std::sqrt(fsky_ratio);
    for (int p = 0; p < X.np; ++p)
        for (int l = 0; l < nlq; ++l) {
            const double s1 = X.bl[l] * X.bl[l] * X.cl[l] + X.nl[l];
            const double s2 = rf * X.bte[l] * X.bte[l] * X.cte[l] + X.nte[l];
            const size_t o = ((size_t)p * nlq + l) * X.nx;
            for (int x = 0; x < X.nx; ++x) qc[o + x] = s1 * qt_E[o + x] - s2 * X.qt[o + x];
        }
    MCMB_TRY(X.d_qc.alloc(sizeof(double) * nq));
    MCMB_HIP(hipMemcpy(X.d_qc.p, qc.data(), sizeof(double) * nq, hipMemcpyHostToDevice));
    X.pol = true;
    X.pol_generation = next_generation();
    return MCMB_OK;
}

extern "C" int mcmb_qtilde_E_info(mcmb_ctx* c, double* fsky_ratio, long* generation)
{
    if (!c || !c->xs.ready || !c->xs.pol) return fail(MCMB_ERR_STATE, "mcmb_qtilde_E_info: no E-mode q-tilde setup");
    if (fsky_ratio) *fsky_ratio = c->xs.fsky_ratio;
    if (generation) *generation = c->xs.pol_generation;
    return MCMB_OK;
}

extern "C" int mcmb_qtilde_info(mcmb_ctx* c, int* np, int* lmax_q, int* nx, int* nmodes, long* generation)
{
    if (!c || !c->xs.ready) return fail(MCMB_ERR_STATE, "mcmb_qtilde_info: no q-tilde setup");
    if (np) *np = c->xs.np;
    if (lmax_q) *lmax_q = c->xs.lq;
    if (nx) *nx = c->xs.nx;
    if (nmodes) *nmodes = c->xs.n;
    if (generation) *generation = c->xs.generation;
    return MCMB_OK;
}

extern "C" int mcmb_qtilde_xvec(mcmb_ctx* c, double* xvec)
{
    if (!c || !c->xs.ready) return fail(MCMB_ERR_STATE, "mcmb_qtilde_xvec: no q-tilde setup");
    if (!xvec) return fail(MCMB_ERR_ARG, "mcmb_qtilde_xvec: null output");
    std::copy(c->xs.xvec.begin(), c->xs.xvec.end(), xvec);
    return MCMB_OK;
}

static int cross_checks(mcmb_ctx* c, const char* who)
{
    if (!c) return fail(MCMB_ERR_ARG, "%s: null context", who);
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "%s: call mcmb_lspace_set (the late basis) first", who);
    if (!c->xs.ready) return fail(MCMB_ERR_STATE, "%s: call mcmb_qtilde_set first", who);
    return MCMB_OK;
}

// ---- 3D form -------------------------------------------------------------------

extern "C" int mcmb_gamma_cross_3d(mcmb_ctx* c, int rank, int nranks, double* gamma_host, double* gamma_dev)
{
    MCMB_TRY(cross_checks(c, "mcmb_gamma_cross_3d"));
    if (nranks < 1) nranks = 1;
    if (rank < 0 || rank >= nranks) return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_3d: rank %d of %d", rank, nranks);
    MCMB_TRY(set_device(c));
    const LSpace& L = c->ls;
    const CrossSet& X = c->xs;
    const int Lq = X.lq, nlate = L.n, nprim = X.n, nx = X.nx, np = X.np;
    const int tlo = std::max(2, L.lmin), thi = std::min(L.lmax, Lq);
    // rank share of the walk of T(2, Lq), split as decompose_gamma_TTT_3D_mpi
    // splits its point count (calculate_gamma_3d.c:100-136)
    long count = 0;
    for (int i = 2; i <= Lq; ++i)
        for (int j = i; j <= Lq; ++j) {
            const int kb = j + (i & 1), ke = std::min(i + j, Lq);
            if (kb <= ke) count += (ke - kb) / 2 + 1;
        }
    const long blk = count / nranks, left = count - blk * nranks;
    const long p0 = rank * blk + std::min<long>(rank, left), p1 = p0 + blk + (rank < left ? 1 : 0);
    // chunks of the share's points with every leg in [tlo, thi]
    std::vector<Chunk> chunks;
    long pos = 0;
    for (int i = 2; i <= Lq && pos < p1; ++i)
        for (int j = i; j <= Lq && pos < p1; ++j) {
            const int kb = j + (i & 1), ke = std::min(i + j, Lq);
            if (kb > ke) continue;
            const long nk = (ke - kb) / 2 + 1;
            const long a = std::max(pos, p0), e = std::min(pos + nk, p1);
            if (a < e && i >= tlo) {
                int k0 = kb + 2 * (int)(a - pos), k1 = kb + 2 * (int)(e - pos - 1);   // inclusive
                k1 = std::min(k1, thi - ((thi - kb) & 1));
                for (int k = k0; k <= k1; k += 2 * CH)
                    chunks.push_back(Chunk{i, j, k, std::min(CH, (k1 - k) / 2 + 1)});
            }
            pos += nk;
        }
    const size_t gb = sizeof(double) * (size_t)nlate * nprim;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(gb));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_HIP(hipMemsetAsync(dst, 0, gb, c->stream));
    if (!chunks.empty()) {
        // per-leg tables: lw = (2l+1)^(-1/3), s4 = (2l+1)^(1/6) / sqrt(C_l + N_l / b_l^2) (l > 1)
        std::vector<double> lw(L.lmax + 1, 0.0), s4(Lq + 1, 0.0);
        for (int l = 2; l <= L.lmax; ++l) lw[l] = 1.0 / std::pow(2.0 * l + 1.0, 1.0 / 3.0);
        for (int l = 2; l <= Lq; ++l)
            s4[l] = std::pow(2.0 * l + 1.0, 1.0 / 6.0) / std::sqrt(X.cl[l] + X.nl[l] / (X.bl[l] * X.bl[l]));
        std::vector<double> cw;
        MCMB_TRY(spline_weights(X.xvec, 2, cw));
        DevBuf d_lw, d_s4, d_cw, d_q3, d_ch;
        MCMB_TRY(d_lw.alloc(sizeof(double) * lw.size()));
        MCMB_TRY(d_s4.alloc(sizeof(double) * s4.size()));
        MCMB_TRY(d_cw.alloc(sizeof(double) * nx));
        const long nq3 = (long)(Lq + 1) * nx * np;
        MCMB_TRY(d_q3.alloc(sizeof(double) * nq3));
        MCMB_HIP(hipMemcpyAsync(d_lw.p, lw.data(), sizeof(double) * lw.size(), hipMemcpyHostToDevice, c->stream));
        MCMB_HIP(hipMemcpyAsync(d_s4.p, s4.data(), sizeof(double) * s4.size(), hipMemcpyHostToDevice, c->stream));
        MCMB_HIP(hipMemcpyAsync(d_cw.p, cw.data(), sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
        MCMB_TRY(launch_guard("q3", (unsigned long long)((nq3 + 255) / 256), 256));
        hipLaunchKernelGGL(cx_q3, dim3((unsigned)((nq3 + 255) / 256)), dim3(256), 0, c->stream, X.d_qt.as<double>(),
                           d_s4.as<double>(), np, Lq + 1, nx, d_q3.as<double>());
        MCMB_HIP(hipGetLastError());
        // blocks of chunks: At (nlate x K) and B (K x nprim) under 2 GiB each
        const long kmax = std::max<long>(CH, (((long)1 << 31) / (8l * std::max(nlate, nprim))) / CH * CH);
        const long cpb = std::max<long>(1, kmax / CH);
        const long nck = (long)chunks.size();
        const long Kb = std::min(nck, cpb) * CH;
        DevBuf d_At, d_B, d_P;
        MCMB_TRY(d_ch.alloc(sizeof(Chunk) * (size_t)std::min(nck, cpb)));
        MCMB_TRY(d_At.alloc(sizeof(double) * (size_t)nlate * Kb));
        MCMB_TRY(d_B.alloc(sizeof(double) * (size_t)Kb * nprim));
        const int tiles = ((nlate + 127) / 128) * ((nprim + 127) / 128);
        MCMB_HIP(hipEventRecord(c->ev0, c->stream));
        for (long c0 = 0; c0 < nck; c0 += cpb) {
            const long ncb = std::min(cpb, nck - c0), K = ncb * CH;
            MCMB_HIP(hipMemcpyAsync(d_ch.p, chunks.data() + c0, sizeof(Chunk) * ncb, hipMemcpyHostToDevice,
                                    c->stream));
            MCMB_TRY(launch_guard("cross B side", (unsigned long long)ncb, 256));
            hipLaunchKernelGGL(cx3_bside, dim3((unsigned)ncb), dim3(256), 0, c->stream, d_ch.as<Chunk>(),
                               (const double*)d_q3.as<double>(), np, nx, X.d_x.as<double>(), d_cw.as<double>(),
                               X.d_modes.as<int>(), nprim, d_B.as<double>(), nprim);
            hipLaunchKernelGGL(cx3_aside, dim3((unsigned)ncb), dim3(256), 0, c->stream, d_ch.as<Chunk>(),
                               L.d_q.as<double>(), L.ldq, L.d_modes.as<int>(), nlate,
                               L.h2mode == MCMB_H2_ASYMPTOTIC ? 1 : 0, L.d_ftab.as<double>(), d_lw.as<double>(),
                               d_At.as<double>(), K);
            MCMB_HIP(hipGetLastError());
            // split K so that about 1024 workgroups run, each over >= 512 points
            const long ksteps = K / 16;
            int nsplit = (int)std::max<long>(1, std::min<long>((1024 + tiles - 1) / tiles, ksteps / 32));
            const int kper = (int)((ksteps + nsplit - 1) / nsplit);
            nsplit = (int)((ksteps + kper - 1) / kper);
            MCMB_TRY(d_P.alloc(sizeof(double) * (size_t)nsplit * nlate * nprim));
            hipLaunchKernelGGL(cx_gemm, dim3((unsigned)tiles, (unsigned)nsplit), dim3(256), 0, c->stream,
                               (const double*)d_At.as<double>(), K, (const double*)d_B.as<double>(), (long)nprim,
                               nlate, nprim, K, kper, d_P.as<double>());
            const long MN = (long)nlate * nprim;
            hipLaunchKernelGGL(cx_reduce, dim3((unsigned)((MN + 255) / 256)), dim3(256), 0, c->stream,
                               (const double*)d_P.as<double>(), nsplit, MN, 6.0 * X.dphi * X.dphi, 1, dst);
            MCMB_HIP(hipGetLastError());
        }
        MCMB_HIP(hipEventRecord(c->ev1, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        float ms = 0.f;
        MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        c->last_ms = ms;
        c->last_launches = (int)(4 * ((nck + cpb - 1) / cpb));
    }
    if (gamma_host) MCMB_TRY(copy_to_host(c, gamma_host, dst, gb));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

// ---- Gauss-Legendre form ---------------------------------------------------------

// Component pol (0 TTT .. 3 EEE) of the GL cross-Gamma into dst (n_late x
// n_prim): per late field in use, the Legendre table times the field's l
// weight, M[x][p][q][s] over the field's range, then the node and x sums.
static int cross_gl_launch(mcmb_ctx* c, int pol, int spline, int npt, double* dst)
{
    static const int fld[4][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const LSpace& L = c->ls;
    const CrossSet& X = c->xs;
    const PolSet& E = c->es;
    const int* ord = pol == 0 ? L.order.data() : E.order[pol - 1].data();
    const int nlate = pol == 0 ? L.n : E.n, np = X.np, nx = X.nx, nprim = X.n;
    const int lo[2] = {L.lmin, E.lmin}, hi[2] = {L.lmax, E.lmax}, ldb[2] = {L.ldq, E.ldq};
    int R[2] = {0, 0};
    for (int m = 0; m < nlate; ++m)
        for (int b = 0; b < 3; ++b) {
            const int f = fld[pol][b], p = ord[3 * m + b];
            if (p >= ldb[f])
                return fail(MCMB_ERR_ORDER, "GL cross-Gamma: late mode %d slot %d uses column %d of %d", m, b, p, ldb[f]);
            R[f] = std::max(R[f], p + 1);
        }
    for (int f = 0; f < 2; ++f) {
        if (R[f] == 0) continue;
        if (R[f] > XR) return fail(MCMB_ERR_ARG, "GL cross-Gamma: late basis width %d > %d", R[f], XR);
        if (hi[f] > X.lq)
            return fail(MCMB_ERR_ARG, "GL cross-Gamma: late lmax %d beyond the q-tilde multipoles (%d)", hi[f], X.lq);
    }
    std::vector<double> cw;
    const double *mu, *w;
    MCMB_TRY(gl_nodes(c, npt, &mu, &w));
    MCMB_TRY(spline_weights(X.xvec, spline, cw));
    // l weights: TTw = (2l+1)^(5/6) b_TT / sqrt(b_TT^2 C_TT + N_TT), EEw =
    // (2l+1)^(5/6) b_EE / sqrt(s1 (s1 s4 - s3^2)) (calculate_gamma.c:73-84,
    // 143-159), 0 for l <= 1
    std::vector<double> tw[2];
    for (int f = 0; f < 2; ++f) {
        if (R[f] == 0) continue;
        tw[f].assign(hi[f] + 1, 0.0);
        for (int l = std::max(2, lo[f]); l <= hi[f]; ++l) {
            const double s1 = X.bl[l] * X.bl[l] * X.cl[l] + X.nl[l];
            if (f == 0) {
                tw[f][l] = std::pow(2.0 * l + 1.0, 5.0 / 6.0) * (X.bl[l] / std::sqrt(s1));
            } else {
                const double s3 = X.bl[l] * X.bee[l] * X.cte[l] + X.nte[l];
                const double s4 = X.bee[l] * X.bee[l] * X.cee[l] + X.nee[l];
                tw[f][l] = std::pow(2.0 * l + 1.0, 5.0 / 6.0) * (X.bee[l] / std::sqrt(s1 * (s1 * s4 - s3 * s3)));
            }
        }
    }
    DevBuf d_mu, d_w, d_cw, d_tw[2], d_Pt[2], d_M[2];
    MCMB_TRY(d_mu.alloc(sizeof(double) * npt));
    MCMB_TRY(d_w.alloc(sizeof(double) * npt));
    MCMB_TRY(d_cw.alloc(sizeof(double) * nx));
    MCMB_HIP(hipMemcpyAsync(d_mu.p, mu, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_w.p, w, sizeof(double) * npt, hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_cw.p, cw.data(), sizeof(double) * nx, hipMemcpyHostToDevice, c->stream));
    for (int f = 0; f < 2; ++f) {
        if (R[f] == 0) continue;
        MCMB_TRY(d_tw[f].alloc(sizeof(double) * tw[f].size()));
        MCMB_TRY(d_Pt[f].alloc(sizeof(double) * (size_t)(hi[f] + 1) * npt));
        MCMB_TRY(d_M[f].alloc(sizeof(double) * (size_t)nx * np * R[f] * npt));
        MCMB_HIP(hipMemcpyAsync(d_tw[f].p, tw[f].data(), sizeof(double) * tw[f].size(), hipMemcpyHostToDevice,
                                c->stream));
    }
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    const double* bq[2] = {L.d_q.as<double>(), E.d_q.as<double>()};
    const double* qsrc[2] = {X.d_qt.as<double>(), X.d_qc.as<double>()};
    int launches = 0;
    for (int f = 0; f < 2; ++f) {
        if (R[f] == 0) continue;
        hipLaunchKernelGGL(cx_legendre, dim3((npt + 255) / 256), dim3(256), 0, c->stream, d_mu.as<double>(), npt,
                           lo[f], hi[f], d_tw[f].as<double>(), d_Pt[f].as<double>());
        hipLaunchKernelGGL(cx_mgemm, dim3((npt + 63) / 64, (nx + 31) / 32, np * R[f]), dim3(256), 0, c->stream,
                           qsrc[f], X.lq + 1, nx, bq[f], ldb[f], R[f], d_Pt[f].as<double>(), npt, lo[f], hi[f], np,
                           d_M[f].as<double>());
        launches += 2;
    }
    const unsigned tiles = (unsigned)(((nlate + 63) / 64) * ((nprim + 63) / 64));
    const double scale = X.dphi * X.dphi / (8.0 * M_PI);
    const int* ml = pol == 0 ? L.d_modes.as<int>() : E.d_modes[pol - 1].as<int>();
    if (pol == 0 || pol == 3) {
        const int f = pol == 0 ? 0 : 1;
        hipLaunchKernelGGL(cx_gl_gamma, dim3(tiles), dim3(256), 0, c->stream, (const double*)d_M[f].as<double>(), np,
                           R[f], nx, npt, d_w.as<double>(), X.d_x.as<double>(), d_cw.as<double>(), ml, nlate,
                           X.d_modes.as<int>(), nprim, scale, dst);
    } else if (pol == 1) {
        hipLaunchKernelGGL(cx_gl_gamma_mixed<1>, dim3(tiles), dim3(256), 0, c->stream,
                           (const double*)d_M[0].as<double>(), R[0], (const double*)d_M[1].as<double>(), R[1], np, nx,
                           npt, d_w.as<double>(), X.d_x.as<double>(), d_cw.as<double>(), ml, nlate,
                           X.d_modes.as<int>(), nprim, scale, dst);
    } else {
        hipLaunchKernelGGL(cx_gl_gamma_mixed<2>, dim3(tiles), dim3(256), 0, c->stream,
                           (const double*)d_M[0].as<double>(), R[0], (const double*)d_M[1].as<double>(), R[1], np, nx,
                           npt, d_w.as<double>(), X.d_x.as<double>(), d_cw.as<double>(), ml, nlate,
                           X.d_modes.as<int>(), nprim, scale, dst);
    }
    MCMB_HIP(hipGetLastError());
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = launches + 1;
    return MCMB_OK;
}

extern "C" int mcmb_gamma_cross_gl(mcmb_ctx* c, int spline, int npt, double* gamma_host, double* gamma_dev)
{
    MCMB_TRY(cross_checks(c, "mcmb_gamma_cross_gl"));
    if (spline != MCMB_SPLINE_FMM && spline != MCMB_SPLINE_NATURAL)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_gl: spline %d", spline);
    const int lmax = c->ls.lmax;
    if (npt <= 0) npt = 3 * lmax / 2 + 1;
    if (2 * npt - 1 < 3 * lmax)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_gl: npt %d < (3 lmax + 1)/2 is not exact", npt);
    MCMB_TRY(set_device(c));
    const size_t gb = sizeof(double) * (size_t)c->ls.n * c->xs.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(gb));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_TRY(cross_gl_launch(c, 0, spline, npt, dst));
    if (gamma_host) MCMB_TRY(copy_to_host(c, gamma_host, dst, gb));
    return MCMB_OK;
}

extern "C" int mcmb_gamma_cross_gl_pol(mcmb_ctx* c, int pol, int spline, int npt, double* gamma_host,
                                       double* gamma_dev)
{
    MCMB_TRY(cross_checks(c, "mcmb_gamma_cross_gl_pol"));
    if (pol < MCMB_POL_TTT || pol > MCMB_POL_EEE) return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_gl_pol: pol %d", pol);
    if (spline != MCMB_SPLINE_FMM && spline != MCMB_SPLINE_NATURAL)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_gl_pol: spline %d", spline);
    if (pol != MCMB_POL_TTT && !c->es.ready)
        return fail(MCMB_ERR_STATE, "mcmb_gamma_cross_gl_pol: call mcmb_lspace_set_E first");
    if (pol != MCMB_POL_TTT && !c->xs.pol)
        return fail(MCMB_ERR_STATE, "mcmb_gamma_cross_gl_pol: call mcmb_qtilde_set_E first");
    const int lt = c->ls.lmax, le = c->es.ready ? c->es.lmax : lt;
    const int deg = pol == MCMB_POL_TTT ? 3 * lt : pol == MCMB_POL_TTE ? 2 * lt + le
                  : pol == MCMB_POL_TEE ? lt + 2 * le : 3 * le;
    // init_gamma_glint's node count with polarisation: 3 max(T lmax, E lmax)/2 + 1
    if (npt <= 0) npt = 3 * std::max(lt, le) / 2 + 1;
    if (2 * npt - 1 < deg)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_cross_gl_pol: %d nodes are not exact for degree %d", npt, deg);
    MCMB_TRY(set_device(c));
    const size_t gb = sizeof(double) * (size_t)(pol == MCMB_POL_TTT ? c->ls.n : c->es.n) * c->xs.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(gb));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_TRY(cross_gl_launch(c, pol, spline, npt, dst));
    if (gamma_host) MCMB_TRY(copy_to_host(c, gamma_host, dst, gb));
    return MCMB_OK;
}

extern "C" int mcmb_spline_weights(int kind, int n, const double* x, double* c)
{
    if (!x || !c || n < 2) return fail(MCMB_ERR_ARG, "mcmb_spline_weights: bad arguments");
    std::vector<double> xv(x, x + n), cw;
    MCMB_TRY(spline_weights(xv, kind, cw));
    std::copy(cw.begin(), cw.end(), c);
    return MCMB_OK;
}
