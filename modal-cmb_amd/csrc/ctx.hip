// Context, error reporting and l-space setup of libmodalcmb.
#include "mcmb_internal.h"
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>

namespace mcmb {

static thread_local char g_msg[512] = "";

int fail(int status, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
    return status;
}

long next_generation()
{
    static std::atomic<long> counter{0};
    return ++counter;
}

int DevBuf::alloc(size_t b)
{
    if (b <= bytes && p) return MCMB_OK;
    release();
    if (b == 0) return MCMB_OK;
    hipError_t e = hipMalloc(&p, b);
    if (e != hipSuccess) {
        p = nullptr;
        bytes = 0;
        return fail(MCMB_ERR_NOMEM, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
    }
    bytes = b;
    return MCMB_OK;
}

void DevBuf::release()
{
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

int set_device(mcmb_ctx* ctx)
{
    MCMB_HIP(hipSetDevice(ctx->device));
    return MCMB_OK;
}

// Allocate the pinned staging halves and their events on first use.
int stage_ready(mcmb_ctx* c)
{
    if (c->stage) return MCMB_OK;
    MCMB_HIP(hipHostMalloc(&c->stage, 2 * STAGE_BYTES, hipHostMallocDefault));
    for (int k = 0; k < 2; ++k) MCMB_HIP(hipEventCreateWithFlags(&c->stage_ev[k], hipEventDisableTiming));
    return MCMB_OK;
}

int copy_stream_ready(mcmb_ctx* c)
{
    MCMB_TRY(stage_ready(c));
    if (c->cstream) return MCMB_OK;
    for (int h = 0; h < 2; ++h)
        for (int k = 0; k < PSLOTS; ++k) MCMB_HIP(hipEventCreateWithFlags(&c->slot_ev[h][k], hipEventDisableTiming));
    MCMB_HIP(hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    return MCMB_OK;
}

int copy_to_device(mcmb_ctx* c, void* dev, const void* host, size_t bytes)
{
    if (!bytes) return MCMB_OK;
    const char* env = getenv("MCMB_HOST_COPY");
    if (bytes < STAGE_BYTES / 4 || (env && !strcmp(env, "direct"))) {
        MCMB_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        return MCMB_OK;
    }
    MCMB_TRY(stage_ready(c));
    char* const stage = static_cast<char*>(c->stage);
    const char* const src = static_cast<const char*>(host);
    char* const dst = static_cast<char*>(dev);
    const size_t nk = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    const int nth = std::min(8, host_threads());
    for (size_t k = 0; k < nk; ++k) {
        // half k & 1 is free once the DMA of chunk k - 2 has read it
        if (k >= 2) MCMB_HIP(hipEventSynchronize(c->stage_ev[k & 1]));
        const size_t off = k * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        char* half = stage + (k & 1) * STAGE_BYTES;
        run_threads(nth, [&](int t) {
            const size_t a = len * t / nth, b = len * (t + 1) / nth;
            memcpy(half + a, src + off + a, b - a);
        });
        MCMB_HIP(hipMemcpyAsync(dst + off, half, len, hipMemcpyHostToDevice, c->stream));
        MCMB_HIP(hipEventRecord(c->stage_ev[k & 1], c->stream));
    }
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int copy_to_host(mcmb_ctx* c, void* host, const void* dev, size_t bytes)
{
    if (!bytes) return MCMB_OK;
    const char* env = getenv("MCMB_HOST_COPY");
    if (bytes < 2 * STAGE_BYTES || (env && !strcmp(env, "direct"))) {
        MCMB_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipStreamSynchronize(c->stream));
        return MCMB_OK;
    }
    MCMB_TRY(stage_ready(c));
    char* const stage = static_cast<char*>(c->stage);
    const char* const src = static_cast<const char*>(dev);
    char* const dst = static_cast<char*>(host);
    const size_t nk = (bytes + STAGE_BYTES - 1) / STAGE_BYTES;
    auto issue = [&](size_t k) -> int {
        const size_t off = k * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        MCMB_HIP(hipMemcpyAsync(stage + (k & 1) * STAGE_BYTES, src + off, len, hipMemcpyDeviceToHost, c->stream));
        MCMB_HIP(hipEventRecord(c->stage_ev[k & 1], c->stream));
        return MCMB_OK;
    };
    const int nth = std::min(8, host_threads());
    MCMB_TRY(issue(0));
    for (size_t k = 0; k < nk; ++k) {
        // half (k + 1) & 1 was emptied into host memory at step k - 1
        if (k + 1 < nk) MCMB_TRY(issue(k + 1));
        MCMB_HIP(hipEventSynchronize(c->stage_ev[k & 1]));
        const size_t off = k * STAGE_BYTES, len = std::min(STAGE_BYTES, bytes - off);
        const char* half = stage + (k & 1) * STAGE_BYTES;
        run_threads(nth, [&](int t) {
            const size_t a = len * t / nth, b = len * (t + 1) / nth;
            memcpy(dst + off + a, half + a, b - a);
        });
    }
    return MCMB_OK;
}

// Rows of T(lmin, lmax) in the order get_ijk_next_XXX walks them
// (tetrapyd_tools.c:104-142): i ascending from lmin, j from i up to
// lmax - i%2, k from j + i%2 in steps of 2 up to min(i + j, lmax).  No row is
// empty (j + i%2 <= i + j and <= lmax), so row (i, j) is at roff[i - lmin] +
// j - i, roff holding the running sum of the lmax - i%2 - i + 1 rows of each
// i; host threads fill ranges of i.
void build_lspace_rows(int lmin, int lmax, std::vector<RowDesc>& rows, long& npoints, std::vector<long>* roff_out)
{
    const int ni = std::max(0, lmax - lmin + 1);
    std::vector<long> roff(ni + 1, 0);
    for (int a = 0; a < ni; ++a) {
        const int i = lmin + a, par = i & 1;
        roff[a + 1] = roff[a] + std::max(0, lmax - par - i + 1);
    }
    rows.resize(roff[ni]);
    const int nth = std::max(1, std::min(host_threads(), ni / 64 + 1));
    std::vector<long> pts(nth, 0);
    run_threads(nth, [&](int t) {
        // i ranges balanced by rows
        const long r0 = roff[ni] * t / nth, r1 = roff[ni] * (t + 1) / nth;
        const int a0 = (int)(std::upper_bound(roff.begin(), roff.end(), r0) - roff.begin()) - 1;
        long np = 0;
        for (int a = std::max(0, a0); a < ni && roff[a] < r1; ++a) {
            if (roff[a] < r0) continue;   // this i belongs to the previous thread
            const int i = lmin + a, par = i & 1;
            long r = roff[a];
            for (int j = i; j <= lmax - par; ++j, ++r) {
                const int kb = j + par, ke = std::min(i + j, lmax);
                const int nk = (ke - kb) / 2 + 1;
                rows[r] = RowDesc{i, j, kb, nk};
                np += nk;
            }
        }
        pts[t] = np;
    });
    npoints = 0;
    for (long v : pts) npoints += v;
    if (roff_out) roff_out->swap(roff);
}

// The same rows on the device, from the per-i offsets: one workgroup per i.
__global__ void lspace_rows_kernel(const long* __restrict__ roff, int lmin, int lmax, RowDesc* __restrict__ rows)
{
    const int a = blockIdx.x, i = lmin + a, par = i & 1;
    const long r0 = roff[a];
    for (int j = i + threadIdx.x; j <= lmax - par; j += blockDim.x) {
        const int kb = j + par, ke = min(i + j, lmax);
        rows[r0 + (j - i)] = RowDesc{i, j, kb, (ke - kb) / 2 + 1};
    }
}

// Copy rows (unless rows_on_device: the caller has written d_rows), basis
// and mode table of a point set to the device.
int upload_pointset(mcmb_ctx* c, PointSet& L, const std::vector<int>& order, bool rows_on_device)
{
    std::vector<int> modes(L.n);
    for (int m = 0; m < L.n; ++m) modes[m] = order[3 * m] | (order[3 * m + 1] << 8) | (order[3 * m + 2] << 16);
    if (!rows_on_device) MCMB_TRY(L.d_rows.alloc(sizeof(RowDesc) * L.rows.size()));
    MCMB_TRY(L.d_q.alloc(sizeof(double) * L.q.size()));
    MCMB_TRY(L.d_modes.alloc(sizeof(int) * modes.size()));
    if (!rows_on_device)
        MCMB_HIP(hipMemcpyAsync(L.d_rows.p, L.rows.data(), sizeof(RowDesc) * L.rows.size(), hipMemcpyHostToDevice,
                                c->stream));
    MCMB_HIP(hipMemcpyAsync(L.d_q.p, L.q.data(), sizeof(double) * L.q.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(L.d_modes.p, modes.data(), sizeof(int) * modes.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

}  // namespace mcmb

using namespace mcmb;

extern "C" {

const char* mcmb_last_error(void) { return g_msg; }

const char* mcmb_status_string(int s)
{
    switch (s) {
    case MCMB_OK: return "ok";
    case MCMB_ERR_ARG: return "bad argument";
    case MCMB_ERR_HIP: return "HIP runtime error";
    case MCMB_ERR_NOT_SPD: return "matrix not positive definite";
    case MCMB_ERR_ORDER: return "bad mode order table";
    case MCMB_ERR_IO: return "I/O error";
    case MCMB_ERR_NOMEM: return "out of memory";
    case MCMB_ERR_STATE: return "setup call missing";
    default: return "unknown status";
    }
}

int mcmb_create(mcmb_ctx** out, int device)
{
    if (!out) return fail(MCMB_ERR_ARG, "mcmb_create: null output pointer");
    *out = nullptr;
    int ndev = 0;
    MCMB_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(MCMB_ERR_ARG, "mcmb_create: device %d not in [0,%d)", device, ndev);
    mcmb_ctx* c = new mcmb_ctx();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&c->ev0);
    if (e == hipSuccess) e = hipEventCreate(&c->ev1);
    for (int k = 0; k <= MAX_PHASES && e == hipSuccess; ++k) e = hipEventCreate(&c->evp[k]);
    if (e != hipSuccess) {
        mcmb_destroy(c);
        return fail(MCMB_ERR_HIP, "mcmb_create: %s", hipGetErrorString(e));
    }
    // the allocation-free GL Gram's scratch address on this device (device current)
    int st = gl_scratch_setup(device);
    if (st != MCMB_OK) {
        mcmb_destroy(c);
        return st;
    }
    *out = c;
    return MCMB_OK;
}

int mcmb_destroy(mcmb_ctx* c)
{
    if (!c) return MCMB_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (PointSet* P : {&c->ls, &c->ks}) {
        P->d_rows.release(); P->d_q.release(); P->d_qf.release(); P->d_v.release(); P->d_ftab.release();
        P->d_ftp.release(); P->d_vp.release();
        P->d_modes.release(); P->d_kn.release(); P->d_kw.release();
    }
    c->d_W.release();
    c->d_partial.release();
    c->d_gamma_tmp.release();
    c->d_aux.release();
    tridiag_release(c);
    if (c->stage) (void)hipHostFree(c->stage);
    for (int k = 0; k < 2; ++k)
        if (c->stage_ev[k]) (void)hipEventDestroy(c->stage_ev[k]);
    for (int h = 0; h < 2; ++h)
        for (int k = 0; k < PSLOTS; ++k)
            if (c->slot_ev[h][k]) (void)hipEventDestroy(c->slot_ev[h][k]);
    if (c->cstream) (void)hipStreamDestroy(c->cstream);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (int k = 0; k <= MAX_PHASES; ++k)
        if (c->evp[k]) (void)hipEventDestroy(c->evp[k]);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return MCMB_OK;
}

int mcmb_sync(mcmb_ctx* c)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_sync: null context");
    MCMB_TRY(set_device(c));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_lspace_set(mcmb_ctx* c, int lmin, int lmax, int ldq, const double* q, const double* v,
                    int nmodes, const int* order, int h2_mode)
{
    if (!c || !q || !v || !order) return fail(MCMB_ERR_ARG, "mcmb_lspace_set: null argument");
    if (lmin < 0 || lmax < lmin + 1 || lmax > 8190)
        return fail(MCMB_ERR_ARG, "mcmb_lspace_set: need 0 <= lmin < lmax <= 8190 (got %d, %d)", lmin, lmax);
    if (ldq < 1 || ldq > 32) return fail(MCMB_ERR_ARG, "mcmb_lspace_set: ldq %d not in [1,32]", ldq);
    if (nmodes < 1 || nmodes > 65536) return fail(MCMB_ERR_ARG, "mcmb_lspace_set: nmodes %d", nmodes);
    if (h2_mode != MCMB_H2_EXACT && h2_mode != MCMB_H2_ASYMPTOTIC)
        return fail(MCMB_ERR_ARG, "mcmb_lspace_set: h2_mode %d", h2_mode);
    int R = 0;
    for (int m = 0; m < nmodes; ++m)
        for (int a = 0; a < 3; ++a) {
            int p = order[3 * m + a];
            if (p < 0 || p >= ldq)
                return fail(MCMB_ERR_ORDER, "mcmb_lspace_set: order[%d][%d] = %d outside [0,%d)", m, a, p, ldq);
            if (p + 1 > R) R = p + 1;
        }
    MCMB_TRY(set_device(c));
    LSpace& L = c->ls;
    L.ready = false;
    L.kspace = false;
    L.kstep = 2;
    L.lmin = lmin;
    L.lmax = lmax;
    L.ldq = ldq;
    L.n = nmodes;
    L.R = R;
    L.h2mode = h2_mode;
    L.order.assign(order, order + 3 * (size_t)nmodes);
    L.q.assign(q, q + (size_t)(lmax + 1) * ldq);
    L.v.assign(v, v + (size_t)(lmax + 1));
    std::vector<long> roff;
    build_lspace_rows(lmin, lmax, L.rows, L.npoints, &roff);
    {
        // the device copy of the rows expanded on the device from the per-i offsets
        DevBuf d_roff;
        MCMB_TRY(d_roff.alloc(sizeof(long) * roff.size()));
        MCMB_TRY(L.d_rows.alloc(sizeof(RowDesc) * L.rows.size()));
        MCMB_HIP(hipMemcpy(d_roff.p, roff.data(), sizeof(long) * roff.size(), hipMemcpyHostToDevice));
        if (lmax >= lmin)
            hipLaunchKernelGGL(lspace_rows_kernel, dim3(lmax - lmin + 1), dim3(256), 0, c->stream,
                               d_roff.as<long>(), lmin, lmax, L.d_rows.as<RowDesc>());
        MCMB_HIP(hipGetLastError());
        MCMB_HIP(hipStreamSynchronize(c->stream));
    }
    // f(a) = prod_{k<=a}(2k-1)/(2k) in long double, a <= 3 lmax / 2 + 1, and
    // 1/((2a+1) f(a)), interleaved: ftab[2a], ftab[2a+1]
    int amax = 3 * lmax / 2 + 2;
    std::vector<double> ftab(2 * (size_t)(amax + 1));
    long double acc = 1.0L;
    for (int a = 0; a <= amax; ++a) {
        if (a > 0) acc *= (long double)(2 * a - 1) / (long double)(2 * a);
        ftab[2 * a] = (double)acc;
        ftab[2 * a + 1] = (double)(1.0L / ((long double)(2 * a + 1) * acc));
    }
    MCMB_TRY(L.d_ftab.alloc(sizeof(double) * ftab.size()));
    MCMB_TRY(L.d_v.alloc(sizeof(double) * L.v.size()));
    MCMB_HIP(hipMemcpyAsync(L.d_v.p, L.v.data(), sizeof(double) * L.v.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(L.d_ftab.p, ftab.data(), sizeof(double) * ftab.size(), hipMemcpyHostToDevice, c->stream));
    {
        const long fg = ftab_guard(lmax);
        std::vector<double> fp(fg + ftab.size() + TAB_TAIL, 0.0), vp(L.v.size() + TAB_TAIL, 0.0);
        std::copy(ftab.begin(), ftab.end(), fp.begin() + fg);
        std::copy(L.v.begin(), L.v.end(), vp.begin());
        MCMB_TRY(L.d_ftp.alloc(sizeof(double) * fp.size()));
        MCMB_TRY(L.d_vp.alloc(sizeof(double) * vp.size()));
        MCMB_HIP(hipMemcpy(L.d_ftp.p, fp.data(), sizeof(double) * fp.size(), hipMemcpyHostToDevice));
        MCMB_HIP(hipMemcpy(L.d_vp.p, vp.data(), sizeof(double) * vp.size(), hipMemcpyHostToDevice));
    }
    MCMB_TRY(upload_pointset(c, L, L.order, true));
    L.ready = true;
    L.generation = next_generation();
    return MCMB_OK;
}

int mcmb_lspace_info(mcmb_ctx* c, int* lmin, int* lmax, int* nmodes, long* generation)
{
    if (!c || !c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_info: no l-space setup");
    if (lmin) *lmin = c->ls.lmin;
    if (lmax) *lmax = c->ls.lmax;
    if (nmodes) *nmodes = c->ls.n;
    if (generation) *generation = c->ls.generation;
    return MCMB_OK;
}

long mcmb_lspace_npoints(mcmb_ctx* c)
{
    if (!c || !c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_npoints: no l-space setup");
    return c->ls.npoints;
}

int mcmb_last_kernel_ms(mcmb_ctx* c, double* ms, int* launches)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_last_kernel_ms: null context");
    if (ms) *ms = c->last_ms;
    if (launches) *launches = c->last_launches;
    return MCMB_OK;
}

int mcmb_device_alloc(mcmb_ctx* c, size_t bytes, void** dev)
{
    if (!c || !dev) return fail(MCMB_ERR_ARG, "mcmb_device_alloc: null argument");
    *dev = nullptr;
    MCMB_TRY(set_device(c));
    if (bytes == 0) return MCMB_OK;
    hipError_t e = hipMalloc(dev, bytes);
    if (e != hipSuccess) {
        *dev = nullptr;
        return fail(MCMB_ERR_NOMEM, "mcmb_device_alloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return MCMB_OK;
}

int mcmb_device_free(mcmb_ctx* c, void* dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_device_free: null context");
    if (!dev) return MCMB_OK;
    MCMB_TRY(set_device(c));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    MCMB_HIP(hipFree(dev));
    return MCMB_OK;
}

}  // extern "C"

namespace mcmb {
// dst[i] = src[i] over n 16-byte words, a grid-stride loop
__global__ void copy16_kernel(const int4* __restrict__ src, int4* __restrict__ dst, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}
}  // namespace mcmb

extern "C" {

int mcmb_device_copy(mcmb_ctx* c, void* dst, const void* src, size_t bytes)
{
    if (!c || !dst || !src || (bytes & 15) || (((uintptr_t)dst | (uintptr_t)src) & 15))
        return fail(MCMB_ERR_ARG, "mcmb_device_copy: null or unaligned buffer, or size not a multiple of 16");
    MCMB_TRY(set_device(c));
    const size_t n = bytes / 16;
    if (n) {
        const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(copy16_kernel, dim3(blocks), dim3(256), 0, c->stream, (const int4*)src, (int4*)dst, n);
        MCMB_HIP(hipGetLastError());
    }
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_device_to_host(mcmb_ctx* c, void* host, const void* dev, size_t bytes)
{
    if (!c || (bytes && (!host || !dev))) return fail(MCMB_ERR_ARG, "mcmb_device_to_host: null argument");
    MCMB_TRY(set_device(c));
    MCMB_TRY(copy_to_host(c, host, dev, bytes));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_last_phases(mcmb_ctx* c, int cap, double* ms, const char** names)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_last_phases: null context");
    if (cap < 0 || (cap > 0 && (!ms || !names))) return fail(MCMB_ERR_ARG, "mcmb_last_phases: bad buffers");
    for (int k = 0; k < c->nphases && k < cap; ++k) {
        float t = 0.f;
        MCMB_HIP(hipEventElapsedTime(&t, c->evp[k], c->evp[k + 1]));
        ms[k] = t;
        names[k] = c->phase_names[k];
    }
    return c->nphases;
}

int mcmb_last_setup(mcmb_ctx* c, int cap, double* ms, const char** names)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_last_setup: null context");
    if (cap < 0 || (cap > 0 && (!ms || !names))) return fail(MCMB_ERR_ARG, "mcmb_last_setup: bad buffers");
    for (int k = 0; k < c->nsetup && k < cap; ++k) {
        ms[k] = c->setup_ms[k];
        names[k] = c->setup_names[k];
    }
    return c->nsetup;
}

int mcmb_last_basis_rank(mcmb_ctx* c, int* rank)
{
    if (!c || !rank) return fail(MCMB_ERR_ARG, "mcmb_last_basis_rank: null argument");
    *rank = c->last_rank;
    return MCMB_OK;
}

int mcmb_gamma_tet(mcmb_ctx* c, int precision, double* gamma_host, double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_tet: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_tet: call mcmb_lspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_MIXED)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_tet: precision %d", precision);
    MCMB_TRY(set_device(c));
    int n = c->ls.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    MCMB_TRY(gram_tet_launch(c, c->ls, precision, dst, 0, (int)c->ls.rows.size()));
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_gamma_tet_slice(mcmb_ctx* c, int precision, int rank, int nranks, double* gamma_host,
                         double* gamma_dev, double* work_fraction)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_slice: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_tet_slice: call mcmb_lspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_MIXED)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_slice: precision %d", precision);
    if (nranks < 1 || rank < 0 || rank >= nranks)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_slice: rank %d of %d", rank, nranks);
    MCMB_TRY(set_device(c));
    int n = c->ls.n, lo, hi;
    lspace_slice(c->ls, 32, rank, nranks, &lo, &hi, work_fraction);
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    if (lo >= hi) {
        MCMB_HIP(hipMemsetAsync(dst, 0, sizeof(double) * (size_t)n * n, c->stream));
        c->last_ms = 0.0;
        c->last_launches = 0;
    } else {
        MCMB_TRY(gram_tet_launch(c, c->ls, precision, dst, lo, hi));
    }
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

int mcmb_gamma_tet_lrange(mcmb_ctx* c, int precision, int l1_lo, int l1_hi, double* gamma_host,
                          double* gamma_dev)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_lrange: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_gamma_tet_lrange: call mcmb_lspace_set first");
    if (precision != MCMB_FP64 && precision != MCMB_MIXED)
        return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_lrange: precision %d", precision);
    if (l1_hi < l1_lo) return fail(MCMB_ERR_ARG, "mcmb_gamma_tet_lrange: l1 range [%d, %d)", l1_lo, l1_hi);
    MCMB_TRY(set_device(c));
    // rows are in walk order, l1 ascending: [l1_lo, l1_hi) is one contiguous row range
    const std::vector<RowDesc>& rows = c->ls.rows;
    int lo = 0, hi = (int)rows.size();
    while (lo < hi && rows[lo].i < l1_lo) ++lo;
    int e = lo;
    while (e < hi && rows[e].i < l1_hi) ++e;
    hi = e;
    int n = c->ls.n;
    double* dst = gamma_dev;
    if (!dst) {
        MCMB_TRY(c->d_gamma_tmp.alloc(sizeof(double) * (size_t)n * n));
        dst = c->d_gamma_tmp.as<double>();
    }
    if (lo >= hi) {
        MCMB_HIP(hipMemsetAsync(dst, 0, sizeof(double) * (size_t)n * n, c->stream));
        c->last_ms = 0.0;
        c->last_launches = 0;
    } else {
        MCMB_TRY(gram_tet_launch(c, c->ls, precision, dst, lo, hi));
    }
    if (gamma_host)
        MCMB_TRY(copy_to_host(c, gamma_host, dst, sizeof(double) * (size_t)n * n));
    MCMB_HIP(hipStreamSynchronize(c->stream));
    return MCMB_OK;
}

}  // extern "C"

extern "C" int mcmb_lspace_set_E(mcmb_ctx* c, int E_lmin, int E_lmax, int ldq, const double* qE, const double* vE,
                                 int nmodes, const int* order_TTE, const int* order_TEE, const int* order_EEE)
{
    if (!c || !qE || !vE || !order_TTE || !order_TEE || !order_EEE)
        return fail(MCMB_ERR_ARG, "mcmb_lspace_set_E: null argument");
    if (E_lmin < 0 || E_lmax < E_lmin + 1 || E_lmax > 8190)
        return fail(MCMB_ERR_ARG, "mcmb_lspace_set_E: need 0 <= E_lmin < E_lmax <= 8190 (got %d, %d)", E_lmin, E_lmax);
    if (ldq < 1 || ldq > 32) return fail(MCMB_ERR_ARG, "mcmb_lspace_set_E: ldq %d not in [1,32]", ldq);
    if (nmodes < 1 || nmodes > 65536) return fail(MCMB_ERR_ARG, "mcmb_lspace_set_E: nmodes %d", nmodes);
    const int* tabs[3] = {order_TTE, order_TEE, order_EEE};
    // slot field per table: 0 = T (index into the l-space basis, < 32), 1 = E (< ldq)
    static const int field[3][3] = {{0, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    for (int t = 0; t < 3; ++t)
        for (int m = 0; m < nmodes; ++m)
            for (int a = 0; a < 3; ++a) {
                const int p = tabs[t][3 * m + a], lim = field[t][a] ? ldq : 32;
                if (p < 0 || p >= lim)
                    return fail(MCMB_ERR_ORDER, "mcmb_lspace_set_E: table %d, mode %d, slot %d: index %d outside [0,%d)",
                                t, m, a, p, lim);
            }
    MCMB_TRY(set_device(c));
    PolSet& E = c->es;
    E.ready = false;
    E.lmin = E_lmin;
    E.lmax = E_lmax;
    E.ldq = ldq;
    E.n = nmodes;
    E.q.assign(qE, qE + (size_t)(E_lmax + 1) * ldq);
    E.v.assign(vE, vE + (size_t)(E_lmax + 1));
    MCMB_TRY(E.d_q.alloc(sizeof(double) * E.q.size()));
    MCMB_TRY(E.d_v.alloc(sizeof(double) * E.v.size()));
    MCMB_HIP(hipMemcpyAsync(E.d_q.p, E.q.data(), sizeof(double) * E.q.size(), hipMemcpyHostToDevice, c->stream));
    MCMB_HIP(hipMemcpyAsync(E.d_v.p, E.v.data(), sizeof(double) * E.v.size(), hipMemcpyHostToDevice, c->stream));
    std::vector<int> packed[3];
    for (int t = 0; t < 3; ++t) {
        E.order[t].assign(tabs[t], tabs[t] + 3 * (size_t)nmodes);
        packed[t].resize(nmodes);
        for (int m = 0; m < nmodes; ++m)
            packed[t][m] = tabs[t][3 * m] | (tabs[t][3 * m + 1] << 8) | (tabs[t][3 * m + 2] << 16);
        MCMB_TRY(E.d_modes[t].alloc(sizeof(int) * nmodes));
        MCMB_HIP(hipMemcpyAsync(E.d_modes[t].p, packed[t].data(), sizeof(int) * nmodes, hipMemcpyHostToDevice,
                                c->stream));
    }
    MCMB_HIP(hipStreamSynchronize(c->stream));
    E.ready = true;
    E.generation = next_generation();
    return MCMB_OK;
}

extern "C" int mcmb_lspace_clear_E(mcmb_ctx* c)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_lspace_clear_E: null context");
    PolSet& E = c->es;
    E.ready = false;
    E.generation = -1;
    E.q.clear();
    E.v.clear();
    E.d_q.release();
    E.d_v.release();
    for (int t = 0; t < 3; ++t) {
        E.order[t].clear();
        E.d_modes[t].release();
    }
    return MCMB_OK;
}

extern "C" int mcmb_lspace_E_info(mcmb_ctx* c, int* E_lmin, int* E_lmax, int* nmodes, long* generation)
{
    if (!c || !c->es.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_E_info: no E-mode setup");
    if (E_lmin) *E_lmin = c->es.lmin;
    if (E_lmax) *E_lmax = c->es.lmax;
    if (nmodes) *nmodes = c->es.n;
    if (generation) *generation = c->es.generation;
    return MCMB_OK;
}

extern "C" int mcmb_lspace_E_tables(mcmb_ctx* c, const double** q, int* ldq, const int** order_TTE,
                                    const int** order_TEE, const int** order_EEE, int* nmodes, int* E_lmin,
                                    int* E_lmax)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_lspace_E_tables: null context");
    if (!c->es.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_E_tables: call mcmb_lspace_set_E first");
    if (q) *q = c->es.q.data();
    if (ldq) *ldq = c->es.ldq;
    if (order_TTE) *order_TTE = c->es.order[0].data();
    if (order_TEE) *order_TEE = c->es.order[1].data();
    if (order_EEE) *order_EEE = c->es.order[2].data();
    if (nmodes) *nmodes = c->es.n;
    if (E_lmin) *E_lmin = c->es.lmin;
    if (E_lmax) *E_lmax = c->es.lmax;
    return MCMB_OK;
}

extern "C" int mcmb_lspace_tables(mcmb_ctx* c, const double** q, int* ldq, const int** order, int* nmodes,
                                  int* lmin, int* lmax)
{
    if (!c) return fail(MCMB_ERR_ARG, "mcmb_lspace_tables: null context");
    if (!c->ls.ready) return fail(MCMB_ERR_STATE, "mcmb_lspace_tables: call mcmb_lspace_set first");
    if (q) *q = c->ls.q.data();
    if (ldq) *ldq = c->ls.ldq;
    if (order) *order = c->ls.order.data();
    if (nmodes) *nmodes = c->ls.n;
    if (lmin) *lmin = c->ls.lmin;
    if (lmax) *lmax = c->ls.lmax;
    return MCMB_OK;
}
