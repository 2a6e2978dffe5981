// Internal declarations shared by the HIP translation units of libmodalcmb.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>
#include "modal_cmb.h"

namespace mcmb {

// Record a formatted error message for mcmb_last_error() and return status.
int fail(int status, const char* fmt, ...);

#define MCMB_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::mcmb::fail(MCMB_ERR_HIP, "%s failed: %s (%s:%d)", #call,       \
                                hipGetErrorString(e_), __FILE__, __LINE__);         \
    } while (0)

#define MCMB_TRY(call)                                                              \
    do {                                                                            \
        int s_ = (call);                                                            \
        if (s_ != MCMB_OK) return s_;                                               \
    } while (0)

// The dispatch packet holds the grid size as a 32-bit count of work-items per
// dimension: a 1D launch of blocks x threads >= 2^32 would wrap silently.
// Returns MCMB_OK or records the error.
inline int launch_guard(const char* what, unsigned long long blocks, unsigned threads)
{
    if (blocks == 0 || blocks * threads >= (1ull << 32))
        return fail(MCMB_ERR_ARG, "%s: grid of %llu x %u work-items is outside one dispatch (limit 2^32 - 1)", what,
                    blocks, threads);
    return MCMB_OK;
}

// Index of the unordered basis pair (x, y), x <= y < R: x-major, y from x.
__host__ __device__ __forceinline__ int pair_index(int x, int y, int R)
{
    return x * R - x * (x - 1) / 2 + (y - x);
}

// One row of the ordered tetrapyd: fixed (i, j), k = k0 + kstep * t, t < nk.
struct RowDesc {
    int i, j, k0, nk;
};

// Device buffer that frees itself.  Movable, not copyable (a copy would free
// the allocation twice).
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes)
    {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf& operator=(DevBuf&& o) noexcept
    {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    int alloc(size_t b);
    void release();
    ~DevBuf() { release(); }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Work buffers of the tridiagonal eigensolver (tridiag.hip, dc_tridiag.hip),
// kept by the context between calls: about 7 n^2 doubles after a
// decomposition of order n, freed with the context.
struct TrdWork {
    DevBuf A, V, Wp, xb, dd, ee, tt, st, part, rpart, E, Vt, M, perm, info;
    DevBuf Qb, Ab, Ub, zb, up, org, tau, zh;   // divide and conquer
};

// Weight of a tetrapyd point.
enum WeightKind : int { W_LSPACE_EXACT = 0, W_LSPACE_ASYM = 1, W_KSPACE_GL = 2 };

// A process-wide increasing counter: every new or changed point set takes a
// fresh value, so (address, generation) never repeats, also for temporaries.
long next_generation();

// An ordered tetrapyd point set with its 1D basis and mode table:
//  l-space: rows of T(lmin, lmax), k in steps of 2, weight permsix h^2 v v v;
//  k-space: Gauss-Legendre node indices a <= b <= c with k_c <= k_a + k_b,
//           k in steps of 1, weight permsix w_a w_b w_c / (k_a + k_b + k_c).
struct PointSet {
    bool ready = false;
    bool kspace = false;
    long generation = next_generation();
    int lmin = 0, lmax = 0;           // l-space multipole range; k-space: 0, N-1
    int ldq = 0, n = 0, h2mode = 0, R = 0, kstep = 2;
    long npoints = 0;
    std::vector<RowDesc> rows;        // host copy of the row table (walk order)
    std::vector<int> order;           // [n][3]
    std::vector<double> q;            // [(lmax+1)][ldq]
    std::vector<double> v;            // l-space per-leg weight [lmax+1]
    std::vector<double> kn, kw;       // k-space nodes and weights [N]
    bool uniform = false;             // uniform k-grid (calculate_weight weights)
    double ukmax = 0.0;               // uniform grid: k = ukmax * i / lmax
    DevBuf d_rows, d_q, d_qf, d_v, d_ftab, d_modes, d_kn, d_kw;
    // l-space: the f table and v with zero guards (ftab_guard(lmax) doubles
    // before index 0 of the f table, TAB_TAIL after either table's end): the
    // fused l3 contraction reads f indices in [-lmax - 64, 3 lmax + 158) and
    // v indices below lmax + 64
    DevBuf d_ftp, d_vp;
    // the point set whose 1D basis this one shares (nullptr: its own) and that
    // set's generation: a factorisation of the basis is keyed on them, so point
    // sets cut from one setup (ranges of its walk) reuse it
    const PointSet* basis_of = nullptr;
    long basis_gen = -1;
    const void* basis_key() const { return basis_of ? (const void*)basis_of : (const void*)this; }
    long basis_generation() const { return basis_of ? basis_gen : generation; }
};
constexpr int TAB_TAIL = 256;
inline long ftab_guard(int lmax) { return 2l * (lmax + 64); }
typedef PointSet LSpace;

// Device-side tables of one block of the separable Gram (offsets in ints
// into the context's table buffer) and the plan they belong to.
struct SepBlock {
    int row0 = 0, nrows = 0, nruns = 0, ngroups = 0, nt2 = 0, nt3 = 0;
    bool contig2 = true, contig3 = true;
    size_t o_groups = 0, o_ginfo = 0, o_wofs = 0, o_aidx2 = 0, o_aidx3 = 0, o_task2 = 0, o_task3 = 0, o_tile2 = 0,
           o_tile3 = 0;
    long wsize = 0;   // doubles of the staged weights of the block's row groups
    int jspan = 0;    // largest span of the varying leg over the rows of one group
    bool swap = false;   // groups of rows sharing l2 (else sharing l1)
};
struct SepPlan {
    const void* ps = nullptr;   // point set the tables were built for
    long gen = -1;              // its generation
    int row_lo = -1, row_hi = -1;
    int gsz = 0;                // rows per stage-1 group
    int nb = 0;                 // basis columns of the contractions
    int tm = 0;                 // GEMM tile rows (the tile shape follows)
    int ksplit = 0;             // K splits of the l1 contraction
    long block_rows = 0;        // row cap override the blocks were cut with (0: none)
    bool by_l2 = false;         // row groups share l2 (else l1)
    size_t max_tab = 0;         // ints per block in the table buffer
    long max_rows = 0, max_runs = 0, max_wsize = 0;
    std::vector<SepBlock> blocks;
};

// Rank-revealing factorisation of the pair products of one point set, U =
// P Ct (r orthonormal columns P over l), and the tables of the expansion of
// the contracted tensor back to pairs.  Kept per point set.
struct SepComp {
    const void* ps = nullptr;   // point set whose basis was factorised (PointSet::basis_key)
    long gen = -1;              // its generation (PointSet::basis_generation)
    int NP = 0, nl = 0, nlp = 0;
    int r = -1;                 // rank; 0: no saving, the pairs are contracted as they are
    double gap = 0.0;           // last kept / first dropped pivot norm at the cut (r > 0 found)
    int ldu = 0;                // leading dimension the P table was laid out with
    DevBuf pc, pt, ct, h, tab;  // P by columns, P table [l][a], Ct [a][pair], H, expansion tables
    size_t o_task[3] = {}, o_tile[3] = {};
    int nt[3] = {}, gs[3] = {};
};

// Projected primordial basis q-tilde and the late-side spectra of the
// cross-Gamma: qt[p][l][x] by multipole l = 0..lq, sample grid xvec, the
// primordial mode table, C_l, N_l, b_l (l = 0..lq) and delta_phi.
struct CrossSet {
    bool ready = false;
    long generation = -1;
    int np = 0, lq = 0, nx = 0, n = 0;
    double dphi = 0.0;
    std::vector<double> qt, xvec, cl, nl, bl;
    std::vector<int> order;
    DevBuf d_qt, d_x, d_modes;
    // E-mode part (polarised cross-Gamma): the primordial side of M^E, qc =
    // s1 qt_E - s2 qt_T, and the TE / EE spectra; dropped by a new T setup
    bool pol = false;
    long pol_generation = -1;
    double fsky_ratio = 1.0;
    std::vector<double> cte, cee, nte, nee, bte, bee;
    DevBuf d_qc;
};

// E-mode field of the late basis, the second field of the polarised Gram:
// basis q[l][ldq] (row = multipole), per-leg weight v[l], and the mode tables
// of the mixed components in find_perm_late_XXX's slot order: order[0] TTE
// (T, T, E), order[1] TEE (T, E, E), order[2] EEE; T indices refer to the
// l-space setup's basis.
struct PolSet {
    bool ready = false;
    long generation = -1;
    int lmin = 0, lmax = 0, ldq = 0, n = 0;
    std::vector<double> q, v;
    std::vector<int> order[3];
    DevBuf d_q, d_v, d_modes[3];
};

// Per-sample weights c_x of an x rule (MCMB_SPLINE_FMM, MCMB_SPLINE_NATURAL,
// MCMB_SPLINE_TRAPEZIUM) on the increasing grid x.
int spline_weights(const std::vector<double>& x, int kind, std::vector<double>& cw);

constexpr int MAX_PHASES = 10;
// Per-batch slots of the pinned staging used by the projection's transfers.
constexpr int PSLOTS = 8;

// Host threads of the table builds: up to 16, never more than the hardware
// has; MCMB_HOST_THREADS sets the count (1..64).
inline int host_threads()
{
    if (const char* e = getenv("MCMB_HOST_THREADS")) {
        const int n = atoi(e);
        if (n >= 1 && n <= 64) return n;
    }
    const unsigned hw = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(16u, hw ? hw : 1u));
}

// Run f(t) for t < n on n host threads (the calling thread takes t = 0).
template <class F> void run_threads(int n, F f)
{
    std::vector<std::thread> th;
    for (int t = 1; t < n; ++t) th.emplace_back(f, t);
    f(0);
    for (auto& x : th) x.join();
}

}  // namespace mcmb

// Roots i0 <= i < i1 of the Gauss-Legendre rule of order n (quadrature.c).
extern "C" void mcmb_gl_roots(int n, int i0, int i1, double* x, double* w);

struct mcmb_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;
    int last_launches = 0;
    mcmb::PointSet ls;   // late-time multipole tetrapyd
    mcmb::PointSet ks;   // primordial k-space Gauss-Legendre tetrapyd
    mcmb::PointSet ku;   // primordial k-space uniform-grid tetrapyd (prim walk)
    mcmb::CrossSet xs;   // q-tilde of the cross-Gamma
    mcmb::PolSet es;     // E-mode field of the late basis (polarised Gram)
    mcmb::DevBuf d_W;    // full-cube projection weight W[i][j][k] (k-space)
    mcmb::DevBuf d_wrow, d_wtile;   // W's row of each pair (i, j); each W row tile's nonzero K chunks
    mcmb::DevBuf pz, py, ph, pd, psc, pdm;   // projection batch buffers (Z, Y, H, partials, sin / cos, D)
    mcmb::DevBuf pio;   // a projection call's shapes, inputs and results on the device
    // stage-1 products Y = W Z of the last forward projection, per batch of
    // shapes, reused by a backward call on the same shapes
    std::vector<mcmb::DevBuf> ycache;
    std::vector<double> ycache_theta;
    int ycache_prec = -1;
    long ycache_gen = -1;
    int w_ready = 0;
    // scratch for partial Gram tiles
    mcmb::DevBuf d_partial;
    mcmb::DevBuf d_gamma_tmp;
    mcmb::DevBuf d_aux;
    // separable Gram: weights, row pointers, pair products, stage outputs, tables
    mcmb::DevBuf sg_ws, sg_rp, sg_wd, sg_U, sg_M, sg_Y, sg_G, sg_tab;
    mcmb::DevBuf sg_Ys, sg_Gs, sg_off;   // symmetrised Y, symmetrised G, its row offsets
    int sg_off_np = 0;                   // pair count the offsets were built for
    // the l3 stage's basis table with the exact weight's l3 factor, and what it was built from
    mcmb::DevBuf sg_U3;
    const void* sg_u3_src = nullptr;
    long sg_u3_gen = -1;
    int sg_u3_ldu = 0, sg_u3_nl = 0;
    mcmb::SepPlan sep;
    mcmb::SepComp sepc;
    mcmb::SepComp sepc_e;    // the same for the E field's basis (polarised GL Gram)
    mcmb::SepComp sepc_pt;   // the T basis over the columns the mixed polarised modes use
    mcmb::DevBuf gl_tab;     // task and tile tables of the mixed polarised GL expansion
    // per-launch timing of the last call that records phases (event k+1 - event k)
    hipEvent_t evp[mcmb::MAX_PHASES + 1] = {};
    int nphases = 0;
    int last_rank = -1;   // basis columns of the last separable Gram's contractions (0: the pairs)
    const char* phase_names[mcmb::MAX_PHASES] = {};
    // host wall time of the parts of the last Gram call (setup and the whole
    // call), in call order
    // pinned staging of large device -> host copies: two halves of
    // STAGE_BYTES, allocated on first use
    void* stage = nullptr;
    hipEvent_t stage_ev[2] = {};
    // copies that overlap the context stream's work, and the events of the
    // projection's per-batch slots of the staging (made on first use)
    hipStream_t cstream = nullptr;
    hipEvent_t slot_ev[2][mcmb::PSLOTS] = {};
    double setup_ms[2 * mcmb::MAX_PHASES] = {};
    const char* setup_names[2 * mcmb::MAX_PHASES] = {};
    int nsetup = 0;
    std::chrono::steady_clock::time_point setup_call, setup_mark;
    // rocBLAS handle of the tridiagonal eigensolver's divide and conquer
    // step (rocsolver_dstedc), created on first use
    void* rb_handle = nullptr;
    mcmb::TrdWork trd;
    // the GL Gram over the compressed basis: node table and Phi
    mcmb::DevBuf gl_xw;
    // the last Gauss-Legendre rule the Grams asked for (gl_nodes)
    std::vector<double> gl_x, gl_w;
};

namespace mcmb {
// The Gauss-Legendre rule of order n, kept by the context (the last order
// asked for) and computed on host threads; *x, *w stay valid until the next
// call with another order.
int gl_nodes(mcmb_ctx* ctx, int n, const double** x, const double** w);
}  // namespace mcmb

namespace mcmb {
// Host wall time of the parts of a call: setup_begin at its start, then
// setup_part(name) at the end of each part (the time since the previous mark).
inline void setup_begin(mcmb_ctx* c)
{
    c->nsetup = 0;
    c->setup_call = c->setup_mark = std::chrono::steady_clock::now();
}
inline void setup_part(mcmb_ctx* c, const char* name)
{
    const auto t = std::chrono::steady_clock::now();
    if (c->nsetup < 2 * MAX_PHASES) {
        c->setup_names[c->nsetup] = name;
        c->setup_ms[c->nsetup++] = std::chrono::duration<double, std::milli>(t - c->setup_mark).count();
    }
    c->setup_mark = t;
}
inline void setup_end(mcmb_ctx* c)
{
    const auto t = std::chrono::steady_clock::now();
    if (c->nsetup < 2 * MAX_PHASES) {
        c->setup_names[c->nsetup] = "whole call";
        c->setup_ms[c->nsetup++] = std::chrono::duration<double, std::milli>(t - c->setup_call).count();
    }
}
}  // namespace mcmb

namespace mcmb {
// Build the ordered-tetrapyd row table of T(lmin, lmax) in walk order.
void build_lspace_rows(int lmin, int lmax, std::vector<RowDesc>& rows, long& npoints,
                       std::vector<long>* roff = nullptr);
int gram_tet_launch(mcmb_ctx* ctx, PointSet& P, int precision, double* d_gamma, int row_lo, int row_hi);
// contiguous share [row_lo, row_hi) of rank r of nr, balanced by chunk count
void lspace_slice(const PointSet& L, int kc, int rank, int nranks, int* row_lo, int* row_hi, double* frac);
int upload_pointset(mcmb_ctx* ctx, PointSet& P, const std::vector<int>& order, bool rows_on_device = false);
int gram_gl_launch(mcmb_ctx* ctx, int npt, double* d_gamma);
// The GL Gram of the bound l-space setup through the separable Gram's
// compressed pair basis (node table x, w on the host, npt entries).  *used
// stays false, and nothing is computed, when the pair products have no
// compressed basis (or MCMB_GL_PATH=direct): the caller then takes the direct
// node sum.
int gram_gl_compressed(mcmb_ctx* ctx, int npt, const double* x, const double* w, double* d_gamma, bool* used);
// The EEE component of the polarised GL Gram over the compressed E basis
// (its first R columns): as gram_gl_compressed, on the E setup.
int gram_gl_eee_compressed(mcmb_ctx* ctx, int R, int npt, const double* x, const double* w, double* d_gamma,
                           bool* used);
// The TTE (pol 1) and TEE (pol 2) components over the compressed T and E
// bases (RT, RE: the columns their modes use); *used false when either
// basis has no low-rank factorisation (then the direct node sum applies).
int gram_gl_mixed_compressed(mcmb_ctx* ctx, int pol, int RT, int RE, int npt, const double* x, const double* w,
                             double* d_gamma, bool* used);
// Look up this device's scratch array of the allocation-free GL Gram (once per
// device, with the device current).
int gl_scratch_setup(int device);
int set_device(mcmb_ctx* ctx);
// Copy bytes from device memory dev (written by work on the context stream)
// to host memory, returning when the host copy is complete.  Copies of 32 MiB
// and more go through the context's pinned staging, 16 MiB halves: the next
// half is copied from the device while host threads move the previous one,
// which is 3x to 4x the rate of a pageable hipMemcpy.
int copy_to_host(mcmb_ctx* ctx, void* host, const void* dev, size_t bytes);
// The reverse, host -> device (4 MiB and more staged: host threads fill one
// half while the other is copied to the device).
int copy_to_device(mcmb_ctx* ctx, void* dev, const void* host, size_t bytes);
// The context's pinned staging (2 x STAGE_BYTES at ctx->stage), allocated on
// first use.  Callers other than the two copies above may use it between
// their own stream synchronisations.
constexpr size_t STAGE_BYTES = (size_t)16 << 20;
int stage_ready(mcmb_ctx* ctx);
// The staging, the context's copy stream (default priority) and the slot
// events, made on first use.
int copy_stream_ready(mcmb_ctx* ctx);
// Symmetric eigen-decomposition by Householder tridiagonalisation (n up to
// 8000): evals (host) descending, d_V rows the sign-fixed eigenvectors; with
// normalise, of D^-1 G D^-1 (d_norm[i] = sqrt G_ii).  tridiag_release frees
// the context's work buffers (TrdWork) and rocBLAS handle (made only when
// MCMB_TRIDIAG=rocsolver).
bool tridiag_supported(int n);
int tridiag_eigen(mcmb_ctx* c, int n, const double* d_G, int normalise, double* d_norm, double* evals, DevBuf& d_V);
void tridiag_release(mcmb_ctx* c);
// Eigenpairs of the symmetric tridiagonal matrix with diagonal d and
// off-diagonal e (host arrays of n and n - 1) by divide and conquer: w (host)
// ascending, dQ (device, column-major n x n) column k the eigenvector of w[k].
int dc_tridiag(mcmb_ctx* c, int n, const double* d, const double* e, double* dQ, double* w);
// Per-point weights of P (walk order) into ws, scaled = sqrt(w)/6 if nonzero;
// rowptr[r] = first point of row r (built and uploaded unless rowptr_ready).
// Weights by the point set's kind.
int pointset_weights(mcmb_ctx* ctx, PointSet& P, DevBuf& rowptr, DevBuf& ws, int scaled, bool rowptr_ready = false);
// Factored Gram (row moments x pair products): whether it handles this basis
// width and precision, and the launch over rows [row_lo, row_hi).
bool gram_factored_supported(const PointSet& P, int precision);
int gram_factored_launch(mcmb_ctx* ctx, PointSet& P, int precision, double* d_gamma, int row_lo, int row_hi);
// Separable Gram (contractions over l3, then l2, then l1 of the pair products):
// whether it handles this basis width and precision, and the launch.
bool gram_separable_supported(const PointSet& P, int precision);
int gram_separable_launch(mcmb_ctx* ctx, PointSet& P, int precision, double* d_gamma, int row_lo, int row_hi);
}  // namespace mcmb
