// Symmetric tridiagonal eigen-decomposition by divide and conquer (Cuppen's
// method with the deflation of LAPACK dlaed2 and the eigenvectors of Gu and
// Eisenstat), for the eigen-decomposition of ortho_late_mpi.c:768-776 after
// the Householder tridiagonalisation.
//
// Split tree: a range of size s > 1 splits at m = s / 2 with rho = e[lo + m - 1]:
// T = diag(T1', T2') + |rho| v v^T, v = e_{lo+m-1} + sign(rho) e_{lo+m}, so T1's
// last and T2's first diagonal lose |rho|.  Leaves have size 1.  Merges run
// by depth, deepest first, all merges of one depth batched together.
//
// Merge of blocks [lo, lo+n1) and [lo+n1, lo+s) with eigenpairs (D, Q):
// D + 2|rho| z z^T with z = (last row of Q1, sign(rho) first row of Q2) / sqrt 2
// (|z| = 1).  On the host, in ascending order of D: deflation of the entries
// with 2|rho| |z_i| <= tol, and of a pair with close poles by a Givens
// rotation that zeroes one z component (tol = 8 eps max(|D|, 2|rho|)).  On the
// device, one wave per remaining root (dc_roots): lambda_j = d_org + tau in
// the interval of pole j, d_org the nearer end.  Then zhat_i^2 = prod_j (lambda_j -
// d_i) / prod_{j != i} (d_j - d_i) / rho (Gu-Eisenstat, so the vectors are
// orthogonal to working precision however close the roots), the vectors
// u_j = zhat / (d - lambda_j) normalised, and Q_new = Q_kept U on FP64 MFMA.
// Differences lambda_j - d_i are formed as (d_org - d_i) + tau.
//
// Q is column-major (column k = eigenvector k, the rows of the caller's E),
// ping-ponged between two buffers that both start as the identity: every
// merge writes its whole diagonal block, so entries outside the blocks of a
// depth stay zero and leaves keep their unit column.
#include "mcmb_internal.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

namespace mcmb {

namespace {

typedef double fd4 __attribute__((ext_vector_type(4)));

constexpr double DC_EPS = 2.220446049250313e-16;

struct DcMerge {
    int lo, n1, s;       // block rows/columns [lo, lo + s), the first child's size n1
    int koff, K;         // its kept items at [koff, koff + K) of the depth's item arrays
    long aoff, uoff;     // its gathered kept columns (s x K) and its K x K vectors
    double rho;          // 2 |rho| (z normalised)
    double zsum;         // sum of its kept z^2
};

struct DcRot {
    int ca, cb;          // global columns rotated: a <- cc b - ss a, b <- ss b + cc a
    double cc, ss;
};

}  // namespace

__global__ void dc_identity(double* __restrict__ X, int n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < (long)n * n) X[e] = (e / n == e % n) ? 1.0 : 0.0;
}

// zbuf[g] = Q[g][rowsel[g]] (column g, the row its merge's z takes), or 0
__global__ void dc_zgather(const double* __restrict__ Q, int n, const int* __restrict__ rowsel,
                           double* __restrict__ zbuf)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const int r = rowsel[g];
    zbuf[g] = r >= 0 ? Q[(long)g * n + r] : 0.0;
}

// Workgroup (merge, row block): the merge's rotations in order, applied to 64
// rows of its block (one thread per row).  Rotation (a, b) deflates column a,
// so a is final after it; b is the merge's newest kept column and either ends
// its chain or is the next rotation's a.  So b is carried in a register and
// stored when its chain ends, every column is loaded at most once (before any
// store to it), and the loads of the next DC_RPF rotations are issued ahead.
constexpr int DC_RPF = 8;
__global__ __launch_bounds__(64) void dc_rotate(double* __restrict__ Q, int n, const int4* __restrict__ rinfo,
                                                const DcRot* __restrict__ rots)
{
    const int4 ri = rinfo[blockIdx.y];   // lo, s, first rotation, count
    const int r = ri.x + blockIdx.x * 64 + threadIdx.x;
    if (blockIdx.x * 64 >= ri.y) return;
    const bool in = r < ri.x + ri.y;
    const DcRot* const R = rots + ri.z;
    double carry = 0.0;
    int ccol = -1;
    for (int q0 = 0; q0 < ri.w; q0 += DC_RPF) {
        const int nq = min(DC_RPF, ri.w - q0);
        double va[DC_RPF], vb[DC_RPF];
#pragma unroll
        for (int u = 0; u < DC_RPF; ++u)
            if (u < nq && in) {
                va[u] = Q[(long)R[q0 + u].ca * n + r];
                vb[u] = Q[(long)R[q0 + u].cb * n + r];
            }
#pragma unroll
        for (int u = 0; u < DC_RPF; ++u)
            if (u < nq && in) {
                const DcRot G = R[q0 + u];
                double a = va[u];
                if (G.ca == ccol) a = carry;
                else if (ccol >= 0) Q[(long)ccol * n + r] = carry;
                Q[(long)G.ca * n + r] = G.cc * vb[u] - G.ss * a;
                carry = G.ss * vb[u] + G.cc * a;
                ccol = G.cb;
            }
    }
    if (in && ccol >= 0) Q[(long)ccol * n + r] = carry;
}

__device__ __forceinline__ double dc_wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One wave per item j: the root of its merge's secular equation in the
// interval of pole j (poles dd ascending, weights z2, rho > 0; the last root
// in (d_{K-1}, d_{K-1} + rho |z|^2)).  Out: org[j] (local pole index) and
// tau[j], lambda = dd[org] + tau, the origin being the nearer end pole (by
// the sign of f at the midpoint), so tau and every lambda - d_i keep full
// relative accuracy.  The lanes share the sums over the poles.
//
// With delta_i = dd[i] - dd[org] and f / rho = 1 / rho + psi + phi, psi over
// the poles up to sp (sp = j; K - 2 for the last root), phi over the rest,
// each step solves a two-pole model of f (the scheme of LAPACK dlaed4):
// either psi as a + q / (delta_sp - x) and phi as b + s / (delta_{sp+1} - x),
// matched in value and slope at tau (the middle way), or the origin's pole
// with its exact weight and the rest lumped on the other pole (the fixed
// weight method); the two alternate while |f| falls by less than 10x per
// step.  After two such slow steps, and whenever the model root leaves the
// bracket kept by the sign of f, the bracket is bisected, geometrically when
// its ends differ in scale (towards a pole end from eps^2 of the other).
// Stops when f is at its rounding level, the step is within 2 ulps of tau,
// or the bracket is.
__global__ __launch_bounds__(256) void dc_roots(int nitem, const int* __restrict__ mof, const DcMerge* __restrict__ M,
                                                const double* __restrict__ dd, const double* __restrict__ z2,
                                                int* __restrict__ org, double* __restrict__ tau)
{
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (j >= nitem) return;   // the whole wave
    const DcMerge& G = M[mof[j]];
    const int K = G.K, jl = j - G.koff;
    const double* const D = dd + G.koff;
    const double* const Z = z2 + G.koff;
    const double rho = G.rho, irho = 1.0 / rho;
    const bool last = jl == K - 1;
    const double lo = D[jl], hi = last ? D[K - 1] + rho * G.zsum : D[jl + 1];
    const double mid = 0.5 * (lo + hi);
    double fm = 0.0;
    for (int i = lane; i < K; i += 64) fm += Z[i] / (D[i] - mid);
    fm = irho + dc_wave_sum(fm);
    const int o = (!last && fm < 0.0) ? jl + 1 : jl;
    const double dorg = D[o];
    const int sp = last ? K - 2 : jl;
    const double pa = sp >= 0 ? D[sp] - dorg : 0.0, pb = sp + 1 < K ? D[sp + 1] - dorg : 0.0;
    double a = lo - dorg, b = hi - dorg, t = 0.5 * (a + b);
    bool fw = false;
    int slow = 0;
    double fprev = 0.0;
    for (int it = 0; it < 100; ++it) {
        double psi = 0.0, dpsi = 0.0, phi = 0.0, dphi = 0.0;
        for (int i = lane; i < K; i += 64) {
            const double inv = 1.0 / ((D[i] - dorg) - t);
            const double q = Z[i] * inv;
            if (i <= sp) {
                psi += q;
                dpsi += q * inv;
            } else {
                phi += q;
                dphi += q * inv;
            }
        }
        psi = dc_wave_sum(psi);
        dpsi = dc_wave_sum(dpsi);
        phi = dc_wave_sum(phi);
        dphi = dc_wave_sum(dphi);
        const double f = irho + psi + phi;
        if (f == 0.0) break;
        if (fabs(f) <= 8.0 * DC_EPS * (irho + fabs(psi) + fabs(phi))) break;
        if (f > 0.0) b = t;
        else a = t;
        if (b - a <= 2.0 * DC_EPS * fmax(fabs(a), fabs(b))) break;
        if (it > 0 && f * fprev > 0.0 && fabs(f) > 0.1 * fabs(fprev)) {
            fw = !fw;
            ++slow;
        } else {
            slow = 0;
        }
        fprev = f;
        double tn = 0.0;
        bool ok = false;
        if (slow < 2) {
            double e1, e2;
            if (K == 1) {
                e1 = e2 = (pa - t) + Z[0] * rho;
            } else {
                const double da = pa - t, db = pb - t;
                double q1, s1, A;
                if (!fw) {
                    q1 = dpsi * da * da;
                    s1 = dphi * db * db;
                    A = irho + (psi - dpsi * da) + (phi - dphi * db);
                } else {
                    const double io = 1.0 / (-t);
                    const double wo = Z[o] * io * io;
                    if (o == sp) {
                        q1 = Z[o];
                        s1 = (dpsi + dphi - wo) * db * db;
                    } else {
                        s1 = Z[o];
                        q1 = (dpsi + dphi - wo) * da * da;
                    }
                    A = f - q1 / da - s1 / db;
                }
                const double B = A * (da + db) + q1 + s1, C = A * da * db + q1 * db + s1 * da;
                const double sq = sqrt(fmax(0.0, B * B - 4.0 * A * C));
                const double den = B + copysign(sq, B);
                e1 = 2.0 * C / den;
                e2 = den / (2.0 * A);
            }
            if (t + e1 > a && t + e1 < b) {
                tn = t + e1;
                ok = true;
            } else if (t + e2 > a && t + e2 < b) {
                tn = t + e2;
                ok = true;
            }
        } else {
            slow = 0;
        }
        if (!ok) {
            const double fa = fmax(a, b * DC_EPS * DC_EPS), fb = fmax(-b, -a * DC_EPS * DC_EPS);
            if (a >= 0.0 && b > 4.0 * fa) tn = sqrt(fa * b);
            else if (b <= 0.0 && -a > 4.0 * fb) tn = -sqrt(fb * -a);
            else tn = 0.5 * (a + b);
        }
        const bool done = fabs(tn - t) <= 2.0 * DC_EPS * fabs(tn);
        t = tn;
        if (done) break;
    }
    if (lane == 0) {
        org[j] = o;
        tau[j] = t;
    }
}

// zhat_i = sign(z_i) sqrt(|prod_j (lambda_j - d_i) / prod_{j != i} (d_j - d_i) / rho|)
__global__ void dc_zhat(int nitem, const int* __restrict__ mof, const DcMerge* __restrict__ M,
                        const double* __restrict__ dd, const double* __restrict__ zz, const int* __restrict__ org,
                        const double* __restrict__ tau, double* __restrict__ zh)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nitem) return;
    const DcMerge& G = M[mof[i]];
    const int K = G.K, il = i - G.koff;
    const double* const D = dd + G.koff;
    const int* const O = org + G.koff;
    const double* const T = tau + G.koff;
    const double di = D[il];
    double p = ((D[O[il]] - di) + T[il]) / G.rho;
    for (int j = 0; j < K; ++j) {
        if (j == il) continue;
        p *= ((D[O[j]] - di) + T[j]) / (D[j] - di);
    }
    zh[i] = copysign(sqrt(fabs(p)), zz[i]);
}

// One wave per item j: column j of its merge's U (column-major, ld K), u_i =
// zhat_i / ((d_i - d_org) - tau_j), normalised.
__global__ __launch_bounds__(64) void dc_vecs(const int* __restrict__ mof, const DcMerge* __restrict__ M,
                                              const double* __restrict__ dd, const int* __restrict__ org,
                                              const double* __restrict__ tau, const double* __restrict__ zh,
                                              double* __restrict__ U)
{
    const int j = blockIdx.x, lane = threadIdx.x;
    const DcMerge& G = M[mof[j]];
    const int K = G.K, jl = j - G.koff;
    const double* const D = dd + G.koff;
    const double* const Zh = zh + G.koff;
    const double dorg = D[org[j]], t = tau[j];
    double ss = 0.0;
    for (int i = lane; i < K; i += 64) {
        const double u = Zh[i] / ((D[i] - dorg) - t);
        ss += u * u;
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) ss += __shfl_xor(ss, o);
    const double inv = 1.0 / sqrt(ss);
    double* const Uj = U + G.uoff + (long)jl * K;
    for (int i = lane; i < K; i += 64) Uj[i] = (Zh[i] / ((D[i] - dorg) - t)) * inv;
}

// A[aoff + k s + r] = Q[kcol[koff + k]][lo + r]: the kept columns of each merge's block
__global__ void dc_gather(const double* __restrict__ Q, int n, int nmerge, const DcMerge* __restrict__ M,
                          const int* __restrict__ kcol, double* __restrict__ A)
{
    const int m = blockIdx.y;
    if (m >= nmerge) return;
    const DcMerge& G = M[m];
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)G.s * G.K) return;
    const int k = (int)(e / G.s), r = (int)(e % G.s);
    A[G.aoff + e] = Q[(long)kcol[G.koff + k] * n + G.lo + r];
}

// Q_new[ocol[koff + j]][lo + r] = sum_k A[k s + r] U[j K + k], 64 x 64 tiles
// (merge, I, J) on FP64 MFMA, four waves of 32 x 32.
__global__ __launch_bounds__(256) void dc_gemm(const int4* __restrict__ tiles, const DcMerge* __restrict__ M,
                                               const double* __restrict__ A, const double* __restrict__ U,
                                               const int* __restrict__ ocol, double* __restrict__ Qn, int n)
{
    const int4 tl = tiles[blockIdx.x];   // merge, I, J
    const DcMerge& G = M[tl.x];
    const int s = G.s, K = G.K;
    const double* const Am = A + G.aoff;
    const double* const Um = U + G.uoff;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
    const int rb = 64 * tl.y + 32 * (w >> 1), cb = 64 * tl.z + 32 * (w & 1);
    fd4 acc[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) acc[x][y] = (fd4){0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < K; k0 += 4) {
        const int k = k0 + kq;
        const bool kin = k < K;
        double a[2], b[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
            const int r = rb + 16 * x + l16, c = cb + 16 * x + l16;
            a[x] = (kin && r < s) ? Am[(long)k * s + r] : 0.0;
            b[x] = (kin && c < K) ? Um[(long)c * K + k] : 0.0;
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
                if (r < s && c < K) Qn[(long)ocol[G.koff + c] * n + G.lo + r] = acc[x][y][v];
            }
}

// deflated columns: Q_new[dst][lo + r] = Q[src][lo + r], one workgroup per column
__global__ __launch_bounds__(256) void dc_copycols(const double* __restrict__ Q, double* __restrict__ Qn, int n,
                                                   const int4* __restrict__ cp)
{
    const int4 c = cp[blockIdx.x];   // lo, s, src, dst
    for (int r = threadIdx.x; r < c.y; r += 256) Qn[(long)c.w * n + c.x + r] = Q[(long)c.z * n + c.x + r];
}

namespace {

struct DcNode {
    int lo, s;
};

// split tree: nodes by depth; diagonals lose |rho| at each split
void dc_split(int lo, int s, int depth, std::vector<double>& d, const std::vector<double>& e,
              std::vector<std::vector<DcNode>>& bydepth)
{
    if (s <= 1) return;
    const int m = s / 2;
    const double rho = e[lo + m - 1];
    d[lo + m - 1] -= fabs(rho);
    d[lo + m] -= fabs(rho);
    if ((int)bydepth.size() <= depth) bydepth.resize(depth + 1);
    bydepth[depth].push_back({lo, s});
    dc_split(lo, m, depth + 1, d, e, bydepth);
    dc_split(lo + m, s - m, depth + 1, d, e, bydepth);
}

// Host arrays packed into one pinned region at 256-byte aligned offsets and
// sent by one copy to the same offsets of a device region.
struct Pack {
    char* h;
    char* d;
    size_t off = 0, cap;
    Pack(char* host, char* dev, size_t bytes) : h(host), d(dev), cap(bytes) {}
    template <class T> const T* put(const std::vector<T>& v)
    {
        off = (off + 255) & ~(size_t)255;
        const size_t b = sizeof(T) * v.size();
        if (off + b > cap) return nullptr;
        if (b) std::memcpy(h + off, v.data(), b);
        const T* p = reinterpret_cast<const T*>(d + off);
        off += b;
        return p;
    }
};

}  // namespace

// Eigenpairs of the symmetric tridiagonal (d, e) of order n (host vectors,
// e of n - 1): w ascending (host) and the eigenvectors as the columns of
// the column-major n x n dQ (device), dQ[k n + i] = component i of vector k.
//
// Per depth of the split tree the host waits twice: for z (the deflation
// needs it) and for the roots (their order fixes the output columns).  The
// small per-depth arrays go up in two packed copies through the context's
// pinned staging; its halves hold [0, 4 MiB) the row selections of every
// depth, [4, 8) and [8, 12) MiB the two uploads, and 16 MiB on the
// downloads.  Between its synchronisations a region is rewritten only after
// the copies that read it have completed (each wait orders after them).
int dc_tridiag(mcmb_ctx* c, int n, const double* d_in, const double* e_in, double* dQ, double* w_out)
{
    constexpr size_t R_SEL = 0, R_UP1 = (size_t)4 << 20, R_UP2 = (size_t)8 << 20, R_UPEND = (size_t)12 << 20;
    constexpr size_t R_DOWN = STAGE_BYTES;
    std::vector<double> d(d_in, d_in + n), e(e_in, e_in + std::max(0, n - 1)), w(n);
    std::vector<std::vector<DcNode>> bydepth;
    dc_split(0, n, 0, d, e, bydepth);
    std::copy(d.begin(), d.end(), w.begin());
    const int ndepth = (int)bydepth.size();
    if ((size_t)ndepth * n * sizeof(int) > R_UP1 || (size_t)n * (sizeof(double) + sizeof(int) + sizeof(double)) > STAGE_BYTES)
        return fail(MCMB_ERR_ARG, "divide and conquer: n = %d exceeds the staging", n);
    MCMB_TRY(stage_ready(c));
    char* const hs = static_cast<char*>(c->stage);
    const size_t nn = (size_t)n * n;
    TrdWork& W = c->trd;
    DevBuf &Qb = W.Qb, &Ab = W.Ab, &Ub = W.Ub, &zb = W.zb, &dup = W.up, &dorg = W.org, &dtau = W.tau, &dzh = W.zh;
    MCMB_TRY(Qb.alloc(sizeof(double) * nn));
    // a depth's kept columns and vectors take at most n^2 each (sum s K, sum K^2)
    MCMB_TRY(Ab.alloc(sizeof(double) * nn));
    MCMB_TRY(Ub.alloc(sizeof(double) * nn));
    MCMB_TRY(zb.alloc(sizeof(double) * n));
    MCMB_TRY(dorg.alloc(sizeof(int) * n));
    MCMB_TRY(dtau.alloc(sizeof(double) * n));
    MCMB_TRY(dzh.alloc(sizeof(double) * n));
    MCMB_TRY(dup.alloc(R_UPEND));
    char* const du = dup.as<char>();
    // the buffer the last depth writes is dQ
    double* Qs[2] = {dQ, Qb.as<double>()};
    int cur = ndepth & 1;
    hipLaunchKernelGGL(dc_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, Qs[0], n);
    hipLaunchKernelGGL(dc_identity, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, c->stream, Qs[1], n);
    MCMB_HIP(hipGetLastError());
    // z of every merge: column g's entry in row lo + n1 - 1 (first child) or lo + n1 (second)
    int* const rsel_h = reinterpret_cast<int*>(hs + R_SEL);
    for (int depth = 0; depth < ndepth; ++depth) {
        int* const rs = rsel_h + (size_t)depth * n;
        std::fill(rs, rs + n, -1);
        for (const DcNode& N : bydepth[depth]) {
            const int n1 = N.s / 2;
            for (int t = 0; t < N.s; ++t) rs[N.lo + t] = t < n1 ? N.lo + n1 - 1 : N.lo + n1;
        }
    }
    if (ndepth) MCMB_HIP(hipMemcpyAsync(du + R_SEL, rsel_h, (size_t)ndepth * n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    double* const zall = reinterpret_cast<double*>(hs + R_DOWN);
    int* const org = reinterpret_cast<int*>(hs + R_DOWN + sizeof(double) * n);
    double* const tau = reinterpret_cast<double*>(hs + R_DOWN + sizeof(double) * n + sizeof(int) * ((n + 1) & ~1));
    // host work arrays, reused across depths
    std::vector<DcMerge> M;
    std::vector<int> mof, kcol, ocol, idx, col, keep, dfoff;
    std::vector<double> dd, z2, zz, D, Z;
    std::vector<DcRot> rots;
    std::vector<int4> rinfo, cps, tiles;
    std::vector<std::pair<double, int>> defl, all, mrg;   // (value, column), (value, item or -1 - deflated index)
    for (auto* v : {&mof, &kcol, &ocol, &idx, &col, &keep}) v->reserve(n);
    for (auto* v : {&dd, &z2, &zz, &D, &Z}) v->reserve(n);
    defl.reserve(n);
    all.reserve(n);
    mrg.reserve(n);
    rots.reserve(n);
    for (int depth = ndepth - 1; depth >= 0; --depth) {
        const std::vector<DcNode>& nodes = bydepth[depth];
        hipLaunchKernelGGL(dc_zgather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                           (const double*)Qs[cur], n, (const int*)(du + R_SEL) + (size_t)depth * n, zb.as<double>());
        MCMB_HIP(hipGetLastError());
        MCMB_HIP(hipMemcpyAsync(zall, zb.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        // host: sort, deflate, rotations; kept items
        M.clear(); mof.clear(); kcol.clear(); dd.clear(); z2.clear(); zz.clear();
        rots.clear(); rinfo.clear(); defl.clear(); dfoff.clear();
        long aoff = 0, uoff = 0;
        for (size_t q = 0; q < nodes.size(); ++q) {
            const DcNode& N = nodes[q];
            const int lo = N.lo, s = N.s, n1 = s / 2;
            const double rho0 = e[lo + n1 - 1];
            const double sg = rho0 < 0.0 ? -1.0 : 1.0;
            const double rho = 2.0 * fabs(rho0);
            // both children's eigenvalues are ascending: merge them
            idx.resize(s);
            for (int x = 0, y = n1, t = 0; t < s; ++t) idx[t] = (y >= s || (x < n1 && w[lo + x] <= w[lo + y])) ? x++ : y++;
            D.resize(s); Z.resize(s); col.resize(s);
            double dmax = 0.0;
            for (int t = 0; t < s; ++t) {
                const int o = idx[t];
                D[t] = w[lo + o];
                Z[t] = (o < n1 ? 1.0 : sg) * zall[lo + o] * 0.70710678118654752440;
                col[t] = lo + o;
                dmax = std::max(dmax, fabs(D[t]));
            }
            const double tol = 8.0 * DC_EPS * std::max(dmax, rho);
            keep.clear();
            dfoff.push_back((int)defl.size());
            const int rot0 = (int)rots.size();
            for (int t = 0; t < s; ++t) {
                if (rho * fabs(Z[t]) <= tol) {
                    defl.push_back({D[t], col[t]});
                    continue;
                }
                if (!keep.empty()) {
                    const int kp = keep.back();
                    const double r = hypot(Z[kp], Z[t]);
                    const double cc = Z[kp] / r, ss = Z[t] / r;
                    if (fabs((D[t] - D[kp]) * cc * ss) <= tol) {
                        rots.push_back({col[kp], col[t], cc, ss});
                        const double dk = D[kp] * ss * ss + D[t] * cc * cc;
                        D[t] = D[kp] * cc * cc + D[t] * ss * ss;
                        D[kp] = dk;
                        Z[t] = r;
                        Z[kp] = 0.0;
                        keep.pop_back();
                        defl.push_back({D[kp], col[kp]});
                    }
                }
                keep.push_back(t);
            }
            if (rots.size() > (size_t)rot0) rinfo.push_back(make_int4(lo, s, rot0, (int)rots.size() - rot0));
            // kept poles ascending (a rotation leaves its kept pole between the two it
            // mixed, so the order changes at most among equal values): insertion sort
            for (size_t x = 1; x < keep.size(); ++x)
                for (size_t y = x; y > 0 && D[keep[y]] < D[keep[y - 1]]; --y) std::swap(keep[y], keep[y - 1]);
            DcMerge G;
            G.lo = lo;
            G.n1 = n1;
            G.s = s;
            G.koff = (int)dd.size();
            G.K = (int)keep.size();
            G.aoff = aoff;
            G.uoff = uoff;
            G.rho = rho;
            G.zsum = 0.0;
            for (int t : keep) {
                dd.push_back(D[t]);
                zz.push_back(Z[t]);
                z2.push_back(Z[t] * Z[t]);
                G.zsum += Z[t] * Z[t];
                kcol.push_back(col[t]);
                mof.push_back((int)q);
            }
            aoff += (long)s * G.K;
            uoff += (long)G.K * G.K;
            M.push_back(G);
        }
        dfoff.push_back((int)defl.size());
        const int nitem = (int)dd.size();
        Pack up1(hs + R_UP1, du + R_UP1, R_UP2 - R_UP1);
        const DcMerge* const dM = up1.put(M);
        const int4* const drinfo = up1.put(rinfo);
        const DcRot* const drot = up1.put(rots);
        const int* const dmof = up1.put(mof);
        const double* const ddd = up1.put(dd);
        const double* const dz2 = up1.put(z2);
        const double* const dzz = up1.put(zz);
        const int* const dkcol = up1.put(kcol);
        if (!dM || !drinfo || !drot || !dmof || !ddd || !dz2 || !dzz || !dkcol)
            return fail(MCMB_ERR_ARG, "divide and conquer: depth arrays exceed the staging at n = %d", n);
        MCMB_HIP(hipMemcpyAsync(du + R_UP1, hs + R_UP1, up1.off, hipMemcpyHostToDevice, c->stream));
        if (!rots.empty()) {
            int maxs = 0;
            for (const int4& ri : rinfo) maxs = std::max(maxs, ri.y);
            hipLaunchKernelGGL(dc_rotate, dim3((unsigned)((maxs + 63) / 64), (unsigned)rinfo.size()), dim3(64), 0,
                               c->stream, Qs[cur], n, drinfo, drot);
        }
        if (nitem > 0) {
            hipLaunchKernelGGL(dc_roots, dim3((unsigned)((nitem + 3) / 4)), dim3(256), 0, c->stream, nitem, dmof, dM,
                               ddd, dz2, dorg.as<int>(), dtau.as<double>());
            MCMB_HIP(hipGetLastError());
            MCMB_HIP(hipMemcpyAsync(org, dorg.p, sizeof(int) * nitem, hipMemcpyDeviceToHost, c->stream));
            MCMB_HIP(hipMemcpyAsync(tau, dtau.p, sizeof(double) * nitem, hipMemcpyDeviceToHost, c->stream));
            // the vectors need only the roots: queued before the wait
            hipLaunchKernelGGL(dc_zhat, dim3((unsigned)((nitem + 63) / 64)), dim3(64), 0, c->stream, nitem, dmof, dM,
                               ddd, dzz, (const int*)dorg.as<int>(), (const double*)dtau.as<double>(),
                               dzh.as<double>());
            hipLaunchKernelGGL(dc_vecs, dim3((unsigned)nitem), dim3(64), 0, c->stream, dmof, dM, ddd,
                               (const int*)dorg.as<int>(), (const double*)dtau.as<double>(),
                               (const double*)dzh.as<double>(), Ub.as<double>());
            long maxsk = 0;
            for (const DcMerge& G : M) maxsk = std::max(maxsk, (long)G.s * G.K);
            hipLaunchKernelGGL(dc_gather, dim3((unsigned)((maxsk + 255) / 256), (unsigned)M.size()), dim3(256), 0,
                               c->stream, (const double*)Qs[cur], n, (int)M.size(), dM, dkcol, Ab.as<double>());
            MCMB_HIP(hipGetLastError());
        }
        MCMB_HIP(hipStreamSynchronize(c->stream));
        // eigenvalues of each merge ascending: the output column of every kept
        // item and deflated column
        ocol.resize(nitem);
        cps.clear();
        tiles.clear();
        for (size_t q = 0; q < nodes.size(); ++q) {
            const DcMerge& G = M[q];
            // the roots ascend with j; the deflated values nearly ascend (insertion
            // sort); then the two lists merged
            all.clear();
            for (int t = dfoff[q]; t < dfoff[q + 1]; ++t) all.push_back({defl[t].first, -1 - t});
            for (size_t x = 1; x < all.size(); ++x)
                for (size_t y = x; y > 0 && all[y].first < all[y - 1].first; --y) std::swap(all[y], all[y - 1]);
            const int nd = (int)all.size();
            for (int j = 0; j < G.K; ++j) all.push_back({dd[G.koff + org[G.koff + j]] + tau[G.koff + j], j});
            mrg.resize(all.size());
            std::merge(all.begin(), all.begin() + nd, all.begin() + nd, all.end(), mrg.begin(),
                       [](const std::pair<double, int>& x, const std::pair<double, int>& y) { return x.first < y.first; });
            for (int p = 0; p < G.s; ++p) {
                w[G.lo + p] = mrg[p].first;
                const int it = mrg[p].second;
                if (it >= 0) ocol[G.koff + it] = G.lo + p;
                else cps.push_back(make_int4(G.lo, G.s, defl[-1 - it].second, G.lo + p));
            }
            for (int I = 0; I < (G.s + 63) / 64; ++I)
                for (int J = 0; J < (G.K + 63) / 64; ++J) tiles.push_back(make_int4((int)q, I, J, 0));
        }
        Pack up2(hs + R_UP2, du + R_UP2, R_UPEND - R_UP2);
        const int* const docol = up2.put(ocol);
        const int4* const dtile = up2.put(tiles);
        const int4* const dcp = up2.put(cps);
        if (!docol || !dtile || !dcp)
            return fail(MCMB_ERR_ARG, "divide and conquer: depth arrays exceed the staging at n = %d", n);
        MCMB_HIP(hipMemcpyAsync(du + R_UP2, hs + R_UP2, up2.off, hipMemcpyHostToDevice, c->stream));
        double* const Qn = Qs[cur ^ 1];
        if (!tiles.empty())
            hipLaunchKernelGGL(dc_gemm, dim3((unsigned)tiles.size()), dim3(256), 0, c->stream, dtile, dM,
                               (const double*)Ab.as<double>(), (const double*)Ub.as<double>(), docol, Qn, n);
        if (!cps.empty())
            hipLaunchKernelGGL(dc_copycols, dim3((unsigned)cps.size()), dim3(256), 0, c->stream, (const double*)Qs[cur],
                               Qn, n, dcp);
        MCMB_HIP(hipGetLastError());
        cur ^= 1;
    }
    MCMB_HIP(hipStreamSynchronize(c->stream));
    std::copy(w.begin(), w.end(), w_out);
    return MCMB_OK;
}

}  // namespace mcmb
