// Uniform k-grid tetrapyd of calculate_ortho (orthogonalisation.c:14-86) and
// calculate_eigen (calculate_eigen.c:86-145).
//
// Points are the prim walk of get_ijk_next_prim (tetrapyd_tools.c:485-517)
// from (min,min,min): rows (i, j), i in [min, max], j in [i, max], with k in
// [j, min(max, i+j+2)], ending at (max, max, max).  A point's weight is
//   x = calculate_weight(min, max, i, j, k) * dk^3 / (k1 + k2 + k3),
// dk = kmax / max, k_a = dk * index (the origin, where k1+k2+k3 = 0, weighs
// 0), and its mode product is pijk (orthogonalisation.c:279-295) = the same
// symmetrised product the Gram kernels form per row.  calculate_weight
// (calculate_volume.c:12-132) is evaluated on the device: the 3 x 3 x 3
// neighbourhood is classified (-8 outside [min,max], 0 outside the triangle,
// 1 inside), each of the 8 sub-cubes at the centre adds its cell rule by its
// count of inside corners, times the ordered multiplicity.  The cell sums are
// sums of small integers, so the device values are bit-identical to the host
// restatement; the k-sum and the products are kept uncontracted for the same
// reason.
//
// The Gram over the walk (calculate_ortho) reuses the FP64 / mixed tetrapyd
// Gram kernels on this point set.  The projection (calculate_eigen) is a
// per-mode reduction: one thread per mode, 256 modes per workgroup, chunks of
// 16 points of one row staged in LDS with the weight x shape of up to 8
// shapes; each thread forms its mode product once per point and accumulates
// it for every shape, per point segment, and the segment partials are summed
// in a fixed order.
#include "mcmb_internal.h"
#include <algorithm>
#include <cmath>
#include <vector>

namespace mcmb {

constexpr int UKC = 16;      // points per chunk
constexpr int UNB = 8;       // shapes per pass
constexpr int UQW = 32;      // basis width staged (ldq <= 32)

static void build_prim_rows(int min, int max, std::vector<RowDesc>& rows, long& np)
{
    rows.clear();
    np = 0;
    for (int i = min; i <= max; ++i)
        for (int j = i; j <= max; ++j) {
            const int khi = std::min(max, i + j + 2);
            rows.push_back(RowDesc{i, j, j, khi - j + 1});
            np += khi - j + 1;
        }
}

// Rows of the walk points [p0, p1) (first and last row clipped).
static void clip_rows(const std::vector<RowDesc>& all, long p0, long p1, std::vector<RowDesc>& out)
{
    out.clear();
    long p = 0;
    for (const RowDesc& rd : all) {
        const long a = std::max(p0, p), b = std::min(p1, p + rd.nk);
        if (a < b) out.push_back(RowDesc{rd.i, rd.j, rd.k0 + (int)(a - p), (int)(b - a)});
        p += rd.nk;
        if (p >= p1) break;
    }
}

__device__ __forceinline__ double cell1(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (11 * (a + b + c + e) + 17 * (d + g + f) + 25 * h) / 240.0; }
__device__ __forceinline__ double cell2(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (47 * a + 19 * (b + c + e) + 5 * (d + f + g) + h) / 720.0; }
__device__ __forceinline__ double cell3(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (40 * (b + c) + 35 * (a + d) + 26 * (f + g) + 19 * (e + h)) / 360.0; }
__device__ __forceinline__ double cell4(double a, double b, double c, double d, double e, double f, double g, double h)
{ return (89 * a + 85 * (b + c + e) + 71 * (d + f + g) + 43 * h) / 720.0; }

__device__ double uniform_cell_weight(int min, int max, int i, int j, int k)
{
    int grid[3][3][3];
    int inside = 0;
    for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m)
            for (int n = 0; n < 3; ++n) {
                const int a = i + l - 1, b = j + m - 1, c = k + n - 1;
                int v;
                if (a < min || a > max || b < min || b > max || c < min || c > max) v = -8;
                else if (a > b + c || b > a + c || c > a + b) v = 0;
                else { v = 1; ++inside; }
                grid[l][m][n] = v;
            }
    double w = 0.0;
    if (inside == 27) {
        w = 1.0;
    } else {
        // corner values: only the centre of the 3 x 3 x 3 block is 1
        for (int s1 = 0; s1 < 2; ++s1)
            for (int s2 = 0; s2 < 2; ++s2)
                for (int s3 = 0; s3 < 2; ++s3) {
                    int sum = 0;
                    for (int l = 0; l < 2; ++l)
                        for (int m = 0; m < 2; ++m)
                            for (int n = 0; n < 2; ++n) sum += grid[l + s1][m + s2][n + s3];
                    // pt[x][y][z] = (x,y,z) == (1,1,1) for the sub-cube corners
                    const double a = (s1 == 1 && s2 == 1 && s3 == 1), b = (s1 == 0 && s2 == 1 && s3 == 1);
                    const double c = (s1 == 1 && s2 == 0 && s3 == 1), d = (s1 == 0 && s2 == 0 && s3 == 1);
                    const double e = (s1 == 1 && s2 == 1 && s3 == 0), f = (s1 == 0 && s2 == 1 && s3 == 0);
                    const double g = (s1 == 1 && s2 == 0 && s3 == 0), h = (s1 == 0 && s2 == 0 && s3 == 0);
                    if (sum == 4) {
                        if (grid[1 + s1][0 + s2][1 + s3] == 1) w += cell2(f, e, h, g, b, a, d, c);
                        if (grid[1 + s1][1 + s2][0 + s3] == 1) w += cell2(d, c, b, a, h, g, f, e);
                        if (grid[0 + s1][1 + s2][1 + s3] == 1) w += cell2(g, h, e, f, c, d, a, b);
                    } else if (sum == 5) {
                        w += cell1(a, b, c, d, e, f, g, h);
                    } else if (sum == 6) {
                        if (grid[0 + s1][0 + s2][1 + s3] == 1) w += cell3(g, h, e, f, c, d, a, b);
                        if (grid[1 + s1][0 + s2][0 + s3] == 1) w += cell3(f, h, b, d, e, g, a, c);
                        if (grid[0 + s1][1 + s2][0 + s3] == 1) w += cell3(d, h, c, g, b, f, a, e);
                    } else if (sum == 7) {
                        if (grid[0 + s1][1 + s2][0 + s3] == 0) w += cell4(f, e, h, g, b, a, d, c);
                        if (grid[0 + s1][0 + s2][1 + s3] == 0) w += cell4(d, c, b, a, h, g, f, e);
                        if (grid[1 + s1][0 + s2][0 + s3] == 0) w += cell4(g, h, e, f, c, d, a, b);
                    } else if (sum == 8) {
                        w += 1.0 / 8.0;
                    }
                }
    }
    if (i != k) w *= (i == j || j == k) ? 3.0 : 6.0;
    return w;
}

__device__ __forceinline__ double uniform_point_weight(int min, int max, int i, int j, int k, double factor,
                                                       double cube)
{
#pragma clang fp contract(off)
    double x = uniform_cell_weight(min, max, i, j, k) * cube;
    const double k1 = factor * i, k2 = factor * j, k3 = factor * k;
    double ks = k1 + k2 + k3;
    if (ks > 1e-10) ks = 1.0 / ks;
    return x * ks;
}

// One wave per row: x_p (scaled == 0) or sqrt(x_p)/6 (scaled == 1).
__global__ __launch_bounds__(256) void uniform_weights_kernel(const RowDesc* __restrict__ rows,
                                                              const long* __restrict__ rowptr, long nrows, int min,
                                                              int max, double factor, double cube, int scaled,
                                                              double* __restrict__ out)
{
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= nrows) return;
    const RowDesc rd = rows[row];
    const long base = rowptr[row];
    for (int t = threadIdx.x & 63; t < rd.nk; t += 64) {
        const double w = uniform_point_weight(min, max, rd.i, rd.j, rd.k0 + t, factor, cube);
        out[base + t] = scaled ? sqrt(w) * (1.0 / 6.0) : w;
    }
}

int uniform_weights_launch(mcmb_ctx* c, PointSet& L, const long* rowptr, double* ws, int scaled)
{
    const long nrows = (long)L.rows.size();
    const double factor = L.ukmax * (1.0 / (double)L.lmax);
    const double cube = factor * factor * factor;
    MCMB_TRY(launch_guard("uniform weights", (unsigned long long)((nrows + 3) / 4), 256));
    hipLaunchKernelGGL(uniform_weights_kernel, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, c->stream,
                       L.d_rows.as<RowDesc>(), rowptr, nrows, L.lmin, L.lmax, factor, cube, scaled, ws);
    MCMB_HIP(hipGetLastError());
    return MCMB_OK;
}

// alpha partials: grid (segment, mode tile); thread = mode.
__global__ __launch_bounds__(256) void uniform_project_kernel(
    const RowDesc* __restrict__ rows, const long* __restrict__ rowptr, const int2* __restrict__ seg_start,
    long nchunks_total, int segch, const int* __restrict__ modes, int n, const double* __restrict__ q, int ldq,
    const double* __restrict__ ws, const double2* __restrict__ shapes, int nb, double factor,
    double* __restrict__ partial)
{
    __shared__ double qk[UQW][UKC];
    __shared__ double fw[UNB][UKC];
    __shared__ double qi[UQW], qj[UQW];
    const int tid = threadIdx.x;
    const long seg = blockIdx.x;
    const int m = blockIdx.y * 256 + tid;
    const bool ok = m < n;
    const int pk = ok ? modes[m] : 0;
    const int r = pk & 255, s = (pk >> 8) & 255, t = (pk >> 16) & 255;
    double acc[UNB];
#pragma unroll
    for (int b = 0; b < UNB; ++b) acc[b] = 0.0;
    const long rest = nchunks_total - seg * segch;
    const int nch = rest < segch ? (int)rest : segch;
    int prow = seg_start[seg].x, pc0 = seg_start[seg].y;
    for (int ch = 0; ch < nch; ++ch) {
        const RowDesc rd = rows[prow];
        const int nk = min(UKC, rd.nk - pc0);
        const long pbase = rowptr[prow] + pc0;
        // stage the chunk: basis rows at k, q(i), q(j), weight x shape / 6
        for (int e = tid; e < UQW * UKC; e += 256) {
            const int cc = e / UKC, kk = e % UKC;
            qk[cc][kk] = (cc < ldq && kk < nk) ? q[(long)(rd.k0 + pc0 + kk) * ldq + cc] : 0.0;
        }
        if (tid < UQW) qi[tid] = tid < ldq ? q[(long)rd.i * ldq + tid] : 0.0;
        else if (tid < 2 * UQW) qj[tid - UQW] = (tid - UQW) < ldq ? q[(long)rd.j * ldq + tid - UQW] : 0.0;
        if (tid >= 128 && tid < 128 + UNB * UKC) {
            const int b = (tid - 128) / UKC, kk = (tid - 128) % UKC;
            double v = 0.0;
            if (b < nb && kk < nk) {
#pragma clang fp contract(off)
                const double k1 = factor * rd.i, k2 = factor * rd.j, k3 = factor * (rd.k0 + pc0 + kk);
                const double p111 = k1 * k2 * k3, p222 = (k1 * k1) * (k2 * k2) * (k3 * k3);
                const double fac = p111 > 1e-15 ? 1.0 / (p111 * p111) : 0.0;
                const double2 th = shapes[b];
                const double sh = p222 * (fac * sin(th.x * (k1 + k2 + k3) + th.y));
                v = ws[pbase + kk] * sh * (1.0 / 6.0);
            }
            fw[b][kk] = v;
        }
        __syncthreads();
        const double Srs = ok ? qi[r] * qj[s] + qi[s] * qj[r] : 0.0;
        const double Srt = ok ? qi[r] * qj[t] + qi[t] * qj[r] : 0.0;
        const double Sst = ok ? qi[s] * qj[t] + qi[t] * qj[s] : 0.0;
        for (int kk = 0; kk < nk; ++kk) {
            const double Q6 = Srs * qk[t][kk] + Srt * qk[s][kk] + Sst * qk[r][kk];
#pragma unroll
            for (int b = 0; b < UNB; ++b) acc[b] += fw[b][kk] * Q6;
        }
        __syncthreads();
        pc0 += UKC;
        if (pc0 >= rd.nk) { ++prow; pc0 = 0; }
    }
    if (ok)
        for (int b = 0; b < nb; ++b) partial[(seg * nb + b) * (long)n + m] = acc[b];
}

__global__ void uniform_project_reduce(const double* __restrict__ partial, long nseg, int nb, int n,
                                       double* __restrict__ alpha)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)nb * n) return;
    double s = 0.0;
    for (long g = 0; g < nseg; ++g) s += partial[g * nb * (long)n + e];
    alpha[e] = s;
}

static int check_range(const PointSet& U, long* p0, long* p1, const char* who)
{
    if (*p1 < 0) *p1 = U.npoints;
    if (*p0 < 0 || *p0 > *p1 || *p1 > U.npoints)
        return fail(MCMB_ERR_ARG, "%s: point range [%ld, %ld) outside the walk of %ld points", who, *p0, *p1,
                    U.npoints);
    return MCMB_OK;
}

}  // namespace mcmb

using namespace mcmb;

extern "C" int mcmb_kuniform_set(mcmb_ctx* c, int min, int max, double kmax, int ldq, const double* q, int nmodes,
                                 const int* order)
{
    if (!c || !q || !order) return fail(MCMB_ERR_ARG, "mcmb_kuniform_set: null argument");
    if (min < 0 || max < 1 || min >= max || max > 8192)
        return fail(MCMB_ERR_ARG, "mcmb_kuniform_set: need 0 <= min < max <= 8192 (got %d, %d)", min, max);
    if (!(kmax > 0.0)) return fail(MCMB_ERR_ARG, "mcmb_kuniform_set: kmax must be positive");
    if (ldq < 1 || ldq > UQW) return fail(MCMB_ERR_ARG, "mcmb_kuniform_set: ldq %d not in [1,%d]", ldq, UQW);
    if (nmodes < 1 || nmodes > 65536) return fail(MCMB_ERR_ARG, "mcmb_kuniform_set: nmodes %d", nmodes);
    int R = 0;
    for (int m = 0; m < nmodes; ++m)
        for (int a = 0; a < 3; ++a) {
            const int p = order[3 * m + a];
            if (p < 0 || p >= ldq)
                return fail(MCMB_ERR_ORDER, "mcmb_kuniform_set: order[%d][%d] = %d outside [0,%d)", m, a, p, ldq);
            R = std::max(R, p + 1);
        }
    MCMB_TRY(set_device(c));
    PointSet& U = c->ku;
    U.ready = false;
    U.kspace = true;
    U.uniform = true;
    U.kstep = 1;
    U.lmin = min;
    U.lmax = max;
    U.ukmax = kmax;
    U.ldq = ldq;
    U.n = nmodes;
    U.R = R;
    U.order.assign(order, order + 3 * (size_t)nmodes);
    U.q.assign(q, q + (size_t)(max + 1) * ldq);
    build_prim_rows(min, max, U.rows, U.npoints);
    MCMB_TRY(upload_pointset(c, U, U.order));
    U.ready = true;
    U.generation = next_generation();
    return MCMB_OK;
}

extern "C" int mcmb_kuniform_info(mcmb_ctx* c, int* min, int* max, int* nmodes, double* kmax, long* generation)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_kuniform_info: null context");
    if (!c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_info: call mcmb_kuniform_set first");
    if (min) *min = c->ku.lmin;
    if (max) *max = c->ku.lmax;
    if (nmodes) *nmodes = c->ku.n;
    if (kmax) *kmax = c->ku.ukmax;
    if (generation) *generation = c->ku.generation;
    return MCMB_OK;
}

extern "C" long mcmb_kuniform_npoints(mcmb_ctx* c)
{
    if (!c || !c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_npoints: call mcmb_kuniform_set first");
    return c->ku.npoints;
}

extern "C" long mcmb_kuniform_weights(mcmb_ctx* c, double* w, long cap)
{
    if (!c || !c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_weights: call mcmb_kuniform_set first");
    MCMB_TRY(set_device(c));
    DevBuf rowptr, ws;
    MCMB_TRY(pointset_weights(c, c->ku, rowptr, ws, 0));
    const long nc = std::min(cap, c->ku.npoints);
    if (w && nc > 0)
        MCMB_HIP(hipMemcpyAsync(w, ws.p, sizeof(double) * nc, hipMemcpyDeviceToHost, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return c->ku.npoints;
}

extern "C" int mcmb_kuniform_gram(mcmb_ctx* c, int precision, long p0, long p1, double* gamma_host,
                                  double* gamma_dev)
{
    if (!c || !c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_gram: call mcmb_kuniform_set first");
    if (precision != MCMB_FP64 && precision != MCMB_MIXED)
        return fail(MCMB_ERR_ARG, "mcmb_kuniform_gram: precision must be MCMB_FP64 or MCMB_MIXED");
    MCMB_TRY(check_range(c->ku, &p0, &p1, "mcmb_kuniform_gram"));
    MCMB_TRY(set_device(c));
    const int n = c->ku.n;
    const size_t gb = sizeof(double) * (size_t)n * n;
    DevBuf tmp;
    double* d_g = gamma_dev;
    if (!d_g) {
        MCMB_TRY(tmp.alloc(gb));
        d_g = tmp.as<double>();
    }
    if (p0 == p1) {
        MCMB_HIP(hipMemsetAsync(d_g, 0, gb, c->stream));
    } else if (p0 == 0 && p1 == c->ku.npoints) {
        MCMB_TRY(gram_tet_launch(c, c->ku, precision, d_g, 0, (int)c->ku.rows.size()));
    } else {
        PointSet R;
        R.kspace = R.uniform = true;
        R.kstep = 1;
        R.lmin = c->ku.lmin;
        R.lmax = c->ku.lmax;
        R.ukmax = c->ku.ukmax;
        R.ldq = c->ku.ldq;
        R.n = n;
        R.R = c->ku.R;
        R.order = c->ku.order;
        R.q = c->ku.q;
        R.basis_of = &c->ku;   // same basis: its factorisation is reused across range calls
        R.basis_gen = c->ku.generation;
        clip_rows(c->ku.rows, p0, p1, R.rows);
        R.npoints = p1 - p0;
        MCMB_TRY(upload_pointset(c, R, R.order));
        R.ready = true;
        MCMB_TRY(gram_tet_launch(c, R, precision, d_g, 0, (int)R.rows.size()));
    }
    if (gamma_host) MCMB_TRY(copy_to_host(c, gamma_host, d_g, gb));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

extern "C" int mcmb_kuniform_project(mcmb_ctx* c, int nshape, const double* theta, long p0, long p1,
                                     double* alpha)
{
    if (!c || !c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_project: call mcmb_kuniform_set first");
    if (nshape < 1 || !theta || !alpha) return fail(MCMB_ERR_ARG, "mcmb_kuniform_project: bad arguments");
    MCMB_TRY(check_range(c->ku, &p0, &p1, "mcmb_kuniform_project"));
    MCMB_TRY(set_device(c));
    PointSet& U = c->ku;
    const int n = U.n;
    std::vector<RowDesc> rows;
    clip_rows(U.rows, p0, p1, rows);
    if (rows.empty()) {
        std::fill(alpha, alpha + (size_t)nshape * n, 0.0);
        return MCMB_OK;
    }
    // weights of the whole walk, then the row table of the range with its
    // offsets into them
    DevBuf rowptr_all, ws;
    MCMB_TRY(pointset_weights(c, U, rowptr_all, ws, 0));
    std::vector<long> rp(rows.size());
    {
        long p = 0;
        size_t ri = 0;
        for (const RowDesc& rd : U.rows) {
            if (ri < rows.size() && rd.i == rows[ri].i && rd.j == rows[ri].j) {
                rp[ri] = p + (rows[ri].k0 - rd.k0);
                ++ri;
            }
            p += rd.nk;
            if (ri == rows.size()) break;
        }
    }
    long nchunks = 0;
    for (const RowDesc& rd : rows) nchunks += (rd.nk + UKC - 1) / UKC;
    const int ntiles = (n + 255) / 256;
    long segch = std::max<long>(1, nchunks * ntiles / 4096);
    segch = std::max<long>(segch, (nchunks + 511) / 512);   // at most 512 segments
    std::vector<int2> seg;
    {
        long cidx = 0;
        for (size_t r = 0; r < rows.size(); ++r)
            for (int c0 = 0; c0 < rows[r].nk; c0 += UKC) {
                if (cidx % segch == 0) seg.push_back(make_int2((int)r, c0));
                ++cidx;
            }
    }
    const long nseg = (long)seg.size();
    DevBuf d_rows, d_rp, d_seg, d_sh, d_part, d_alpha;
    MCMB_TRY(d_rows.alloc(sizeof(RowDesc) * rows.size()));
    MCMB_TRY(d_rp.alloc(sizeof(long) * rp.size()));
    MCMB_TRY(d_seg.alloc(sizeof(int2) * seg.size()));
    MCMB_TRY(d_sh.alloc(sizeof(double2) * UNB));
    MCMB_TRY(d_part.alloc(sizeof(double) * (size_t)nseg * UNB * n));
    MCMB_TRY(d_alpha.alloc(sizeof(double) * (size_t)UNB * n));
    MCMB_HIP(hipMemcpyAsync(d_rows.p, rows.data(), sizeof(RowDesc) * rows.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_rp.p, rp.data(), sizeof(long) * rp.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(d_seg.p, seg.data(), sizeof(int2) * seg.size(), hipMemcpyHostToDevice, c->stream));
    const double factor = U.ukmax * (1.0 / (double)U.lmax);
    MCMB_HIP(hipEventRecord(c->ev0, c->stream));
    for (int b0 = 0; b0 < nshape; b0 += UNB) {
        const int nb = std::min(UNB, nshape - b0);
        std::vector<double2> sh(UNB, make_double2(0.0, 0.0));
        for (int b = 0; b < nb; ++b) sh[b] = make_double2(theta[2 * (b0 + b)], theta[2 * (b0 + b) + 1]);
        MCMB_HIP(hipMemcpyAsync(d_sh.p, sh.data(), sizeof(double2) * UNB, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL(uniform_project_kernel, dim3((unsigned)nseg, (unsigned)ntiles), dim3(256), 0, c->stream,
                           d_rows.as<RowDesc>(), d_rp.as<long>(), d_seg.as<int2>(), nchunks, (int)segch,
                           U.d_modes.as<int>(), n, U.d_q.as<double>(), U.ldq, ws.as<double>(), d_sh.as<double2>(),
                           nb, factor, d_part.as<double>());
        MCMB_HIP(hipGetLastError());
        const long ne = (long)nb * n;
        hipLaunchKernelGGL(uniform_project_reduce, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream,
                           d_part.as<double>(), nseg, nb, n, d_alpha.as<double>());
        MCMB_HIP(hipGetLastError());
        MCMB_HIP(hipMemcpyAsync(alpha + (size_t)b0 * n, d_alpha.p, sizeof(double) * ne, hipMemcpyDeviceToHost,
                                c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));   // the shape buffer is reused
    }
    MCMB_HIP(hipEventRecord(c->ev1, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    MCMB_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    c->last_launches = (nshape + UNB - 1) / UNB;
    return MCMB_OK;
}

extern "C" int mcmb_kuniform_tables(mcmb_ctx* c, const double** q, int* ldq, const int** order, int* nmodes,
                                    int* max)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_kuniform_tables: null context");
    if (!c->ku.ready) return fail(MCMB_ERR_STATE, "mcmb_kuniform_tables: call mcmb_kuniform_set first");
    if (q) *q = c->ku.q.data();
    if (ldq) *ldq = c->ku.ldq;
    if (order) *order = c->ku.order.data();
    if (nmodes) *nmodes = c->ku.n;
    if (max) *max = c->ku.lmax;
    return MCMB_OK;
}
