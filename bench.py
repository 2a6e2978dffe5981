"""Benchmark of the Gamma and shape-projection hot path on one MI355X.

Headline (value): seconds per FP64 tetrapyd Gamma at lmax=2500 with 4960 modes
(r <= s <= t < 30), 653,383,126 tetrapyd points, by the separable contraction
over the compressed pair basis: the 465 pair products q_x q_y span 59
directions (rank found by pivoted Gram-Schmidt on the first call for the
setup, kept with it); row moments over l3 with the exact 3j weights formed in
the kernel, FP64 MFMA contractions over l2 and the symmetrised l1 on the 59
columns, the symmetrised tensor expanded back to pairs, 6-term assembly.  One
step = one full Gamma computed from the resident basis, per-leg weights and
mode table into a device buffer; the step is the C-ABI call mcmb_gamma_tet,
timed on the host around the call.

Reported beside it:
  gamma_kernel_ms     HIP-event time of the call and its phases
  basis_rank          columns of the compressed pair basis (0: pairs)
  gamma_first_call_s  the first call on the setup (the warm-up), which also
                      factorises the basis and builds the host row tables
  fp64_separable_pairs_s
                      the same separable contraction over the 465 pair
                      products themselves (MCMB_SEP_COMPRESS=0), per call
  fp64_gl_s           the Gauss-Legendre factorised FP64 Gram (the reference's
                      default algorithm, calculate_orthol) into host memory,
                      steady (second call); fp64_gl_first_call_s the first,
                      which also computes the 3751-node rule
  fp64_gl_pol         the polarised GL components TTE, TEE, EEE over the
                      compressed T and E bases, per call, beside the direct
                      node sum's time and the difference between the two
  fp64_factored_s     the same FP64 Gamma by row moments x pair products
  fp64_direct_s       the point-by-point FP64 contraction on the GPU (the
                      reference's algorithm), timed on one slice of the rows
                      and extrapolated by its work fraction
  mixed_direct_s      the point-by-point contraction in mixed precision (BF16
                      hi+lo x3, FP32 tiles, FP64 sums), same slice method
  gamma_err           max_mn |G - G_x| / sqrt(G_mm G_nn) between the
                      separable Gram and the factored FP64 Gram (full) and the
                      Gauss-Legendre factorised FP64 Gram (full); between the
                      factored FP64 slice and the direct FP64 and mixed slices
                      (same rows)
  eig / beta          eigenvalues of C = D^-1 G D^-1 and the orthonormalised
                      coefficients beta^R = lambda alpha (alpha = G c, seeded
                      c; lambda from the late-side orthonormalisation with the
                      PCA branch) of the separable Gram against those of the
                      factored Gram and of the GL Gram; prim-side Cholesky
  fp64_gamma_other_workloads
                      the same FP64 Gamma at lmax 500 / 220 modes and lmax
                      2000 / 1540 modes (median of 3 calls)
  projection          256-shape batch (omega, phi) at 300 GL nodes, 4960 modes:
                      forward with tangents and the adjoint, FP32 MFMA with
                      FP64 sums; gradients/s and the error against FP64
  roofline            the l3 contraction (the largest phase) against the FP64
                      dense MFMA peak, algorithmic flops 2 x points x rank
  one_shot            a fresh setup's first Gamma into host memory: setup
                      (tables to the device), the first call (basis
                      factorisation, host row tables, buffers) and the copy
  cholesky_spd        the prim-side orthonormalisation (FP64 Cholesky, SYRK
                      and triangular updates on FP64 MFMA, L^-1) and the
                      beta solve with 4 right-hand sides on an SPD 4960 x 4960
                      matrix: the separable Gram with its diagonal raised by
                      1e-4 of itself (the Gram itself is singular in FP64)
  cpu_baseline        the FP64 C oracle on all host cores running the
                      reference's default algorithm for this Gram, the
                      Gauss-Legendre factorisation of calculate_orthol /
                      calculate_gamma_2d, for the whole Gamma; beside it
                      (direct_extrapolated) the direct tetrapyd sum timed on
                      the top i-rows and extrapolated by point count

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 5] [--slices 32]
                       [--cpu-seconds 20] [--no-cpu] [--no-projection] [--no-other-workloads]
                       [--workload lmax2500_r30]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

# numpy's BLAS threads spin for a while after each call and hold the cores
# the library's own host threads (row tables, staged copies) run on: a
# first Gamma right after a numpy matrix product measured 79 ms instead of
# 22 ms, with 1.2 s of CPU time spent in that window.  The bench's numpy
# work is small, so it runs on one BLAS thread (set before numpy loads,
# overriding an inherited thread count).
os.environ["OPENBLAS_NUM_THREADS"] = "1"

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modal-cmb_amd"))

METRIC = "s per Gamma (lmax 2500, 4960 modes) mixed vs FP64 GPU at 1e-10 err; grads/s"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    d[d == 0] = 1.0
    return float((np.abs(A - B) / d).max())


def hbm_traffic(workload, fname="gamma_sep_hbm_traffic.md"):
    """HBM bytes per launch of the dominant kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    p = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(p):
        return None
    for line in open(p):
        m = re.match(r"\s*traffic_bytes_per_launch\[(\w+)\]\s*=\s*([0-9.eE+-]+)", line)
        if m and m.group(1) == workload:
            return float(m.group(2))
    return None


def lspace_rows(L, lmin=2):
    """Rows (l1, l2) of T(lmin, L), as the walk holds them."""
    i = np.arange(lmin, L + 1)
    return int(np.maximum(0, L - (i & 1) - i + 1).sum())


def cpu_baseline_gl(ws, L):
    """The oracle's Gauss-Legendre factorised Gram (the reference's default
    algorithm, calculate_orthol / calculate_gamma_2d: M_ij over 3L/2+1 nodes,
    then the node sum of the 3 x 3 permanents per (m, n)) for the whole
    Gamma on all host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    n = ws["order"].shape[0]
    npt = 3 * L // 2 + 1
    t = time.time()
    O.gamma_gl(2, L, ws["q"], ws["v"], ws["order"], npt)
    dt = time.time() - t
    return {
        "value": round(dt, 3),
        "unit": "s per Gamma",
        "cores": int(O.lib().or_max_threads()),
        "kind": "port",
        "sample": f"the whole Gamma (not a sample): oracle Gauss-Legendre factorised FP64 Gram, {npt} nodes, "
                  f"lmax={L}, n={n} -- the reference's default algorithm (calculate_orthol, "
                  f"orthogonalisation_late.c:239-336; gamma_mpi.c:316 with basis:do3D defaulting to 1)",
    }


def cpu_baseline(ws, L, target_s):
    """Time the oracle's direct tetrapyd Gram over the top i-rows of T(2, L)
    (about target_s seconds of work on all host cores) and scale by points."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    n = ws["order"].shape[0]
    cnt = np.zeros(L + 2, dtype=np.int64)   # points per i-row of T(2, L)
    for i in range(2, L + 1):
        j = np.arange(i, L - (i & 1) + 1)
        ke = np.minimum(i + j, L)
        kb = j + (i & 1)
        cnt[i] = int(np.maximum(0, (ke - kb) // 2 + 1).sum())
    total = int(cnt.sum())
    nthreads = O.lib().or_max_threads()

    def run(i_lo, i_hi):
        t = time.time()
        O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"], 0, i_lo, i_hi)
        return time.time() - t

    def rows_from_top(points, min_rows):
        """Lowest i such that i-rows i..L hold about `points` points (at least min_rows rows)."""
        i, acc = L + 1, 0
        while i > 2 and (acc + cnt[i - 1] <= points or L + 1 - i < min_rows):
            i -= 1
            acc += int(cnt[i])
        return i

    # fixed cost of one call (per-thread partial Grams, their reduction),
    # paid once by a whole Gamma: timed on an empty row range
    t_fixed = run(L + 1, L + 1)
    min_rows = 4 * max(16, nthreads)
    i_lo = rows_from_top(0, min_rows)
    pts = int(cnt[i_lo:].sum())
    t_probe = run(i_lo, L + 1)
    rate = pts / max(t_probe - t_fixed, 1e-3)
    i_lo2 = rows_from_top(int(rate * target_s), min_rows)
    pts2 = int(cnt[i_lo2:].sum())
    t2 = run(i_lo2, L + 1)
    return {
        "value": t_fixed + (t2 - t_fixed) * total / pts2,
        "unit": "s per Gamma (extrapolated)",
        "cores": int(nthreads),
        "kind": "port",
        "sample": f"oracle direct tetrapyd FP64 Gram, i-rows {i_lo2}..{L} of T(2,{L}): {pts2} of {total} points "
                  f"({t2:.1f} s, of which {t_fixed:.1f} s fixed per-call cost), the point-dependent part scaled "
                  f"by point count; n={n}",
    }


def cpu_projection(ks, theta):
    """One shape of the projection batch through the oracle (FP64 direct sum
    over the ordered k-tetrapyd with alpha and both derivatives)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    t = time.time()
    O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], float(theta[0]), float(theta[1]))
    dt = time.time() - t
    return {"value": 1.0 / dt, "unit": "shapes/s (alpha + d alpha/d theta)", "cores": int(O.lib().or_max_threads()),
            "kind": "port", "sample": f"one shape (omega={theta[0]:.3f}, phi={theta[1]:.3f}) of the batch, {dt:.2f} s"}


def bench_other_workloads(modalcmb, W, skip):
    """FP64 separable Gamma of the other l-space workloads, into a
    device buffer: median of 3 calls after one warm-up."""
    out = {}
    for k in sorted(W.WORKLOADS):
        cfg = W.WORKLOADS[k]
        if k == skip or "L" not in cfg:
            continue
        ws = W.lspace(cfg["L"], cfg["R"], cfg["modulated"])
        n = ws["order"].shape[0]
        with modalcmb.Context(0) as ctx:
            ctx.lspace_set(2, cfg["L"], ws["q"], ws["v"], ws["order"])
            d = ctx.device_alloc(n * n * 8)
            ctx.gamma_tet_device(modalcmb.FP64, d)
            ctx.sync()
            ts = []
            for _ in range(3):
                t = time.time()
                ctx.gamma_tet_device(modalcmb.FP64, d)
                ts.append(time.time() - t)
            ctx.device_free(d)
        out[f"lmax{cfg['L']}_modes{n}_s"] = round(float(np.median(ts)), 5)
    return out


def bench_projection(modalcmb, W, args):
    cfg = W.WORKLOADS["gl300_r30_batch256"]
    ks = W.kspace(cfg["N"], cfg["R"])
    theta = W.shape_batch(cfg["batch"], cfg["seed"])
    ctx = modalcmb.Context(0)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    B, n = theta.shape[0], ctx.nk
    gbar = np.random.default_rng(7).standard_normal((B, n))
    for _ in range(2):                        # warm-up (weight cube, batch and product buffers)
        ctx.project_fwd(theta, modalcmb.FP32)
        ctx.project_bwd(theta, gbar, modalcmb.FP32)
    ctx.sync()
    tf, kf, tb, kb = [], [], [], []
    for _ in range(3):
        t = time.time()
        a32, d32 = ctx.project_fwd(theta, modalcmb.FP32)
        tf.append(time.time() - t)
        kf.append(ctx.last_kernel_ms()[0])
        t = time.time()
        g32 = ctx.project_bwd(theta, gbar, modalcmb.FP32)   # reuses the forward's stage-1 products
        tb.append(time.time() - t)
        kb.append(ctx.last_kernel_ms()[0])
    t_fwd, k_fwd, t_bwd, k_bwd = (float(np.median(x)) for x in (tf, kf, tb, kb))
    ctx.project_fwd(theta[::-1].copy(), modalcmb.FP32)  # other shapes: the next adjoint recomputes stage 1
    t = time.time()
    g32b = ctx.project_bwd(theta, gbar, modalcmb.FP32)
    t_bwd_alone = time.time() - t
    assert np.array_equal(g32, g32b)
    a64, d64 = ctx.project_fwd(theta, modalcmb.FP64)
    g64 = ctx.project_bwd(theta, gbar, modalcmb.FP64)
    ctx.sync()
    t = time.time()
    ctx.project_fwd(theta, modalcmb.FP64)       # second FP64 call: its weight cube and buffers exist
    t_fwd64 = time.time() - t
    t = time.time()
    ctx.project_bwd(theta, gbar, modalcmb.FP64)
    t_bwd64 = time.time() - t
    npts = ctx.kspace_npoints()
    ctx.close()

    def rel(x, y):
        return float(np.abs(x - y).max() / np.abs(y).max())

    # stage-1 contraction flops per shape: 2 nnz(W) (2R), nnz(W) the weights
    # inside the triangle over the pair rows i <= j and all k (W is symmetric
    # in i, j; sorted legs a <= b <= c by index, k_c <= k_a + k_b), times the
    # cos and sin column sets per basis index
    N, R = cfg["N"], cfg["R"]
    kn = np.asarray(ks["kn"])
    I, J = np.triu_indices(N)
    kk = np.arange(N)
    lo_ = np.minimum(np.minimum(I[:, None], J[:, None]), kk[None, :])
    hi_ = np.maximum(np.maximum(I[:, None], J[:, None]), kk[None, :])
    mid = I[:, None] + J[:, None] + kk[None, :] - lo_ - hi_
    nnz_w = int((kn[hi_] <= kn[lo_] + kn[mid]).sum())
    del lo_, hi_, mid
    out = {
        "workload": f"{B} shapes S=sin(omega K + phi)/(k1k2k3)^2 form, GL {N}/side (k in [2e-4,0.2]), "
                    f"{n} modes (r<=s<=t<{R}), masked cube {npts} ordered nodes",
        "timing": "median of 3 forward + adjoint rounds after 2 warm-up rounds",
        "fwd_tangent_s": round(t_fwd, 4), "fwd_kernel_ms": round(k_fwd, 2),
        "adjoint_s": round(t_bwd, 4), "adjoint_kernel_ms": round(k_bwd, 2),
        "adjoint_note": "right after the forward on the same shapes, reusing its stage-1 products; "
                        "adjoint_alone_s recomputes them",
        "adjoint_alone_s": round(t_bwd_alone, 4),
        "shapes_per_s_fwd_tangent": round(B / t_fwd, 1),
        "grads_per_s_fwd_tangent": round(2 * B / t_fwd, 1),
        "grads_per_s_fwd_plus_adjoint": round(2 * B / (t_fwd + t_bwd), 1),
        "fp64": {"fwd_tangent_s": round(t_fwd64, 4), "adjoint_s": round(t_bwd64, 4),
                 "grads_per_s_fwd_tangent": round(2 * B / t_fwd64, 1),
                 "grads_per_s_fwd_plus_adjoint": round(2 * B / (t_fwd64 + t_bwd64), 1),
                 "note": "FP64 MFMA k-sum (within 1e-12 of the oracle per shape), one timed call each"},
        "grads_per_s_at_1e-8": round(2 * B / t_fwd64, 1),
        "grads_note": "gradients/s within 1e-8 of FP64 per shape: the FP64 path (the FP32 path, faster, is within "
                      "1e-8 of the batch's largest coefficient but only ~1e-6 of a shape's own for oscillatory "
                      "shapes)",
        "err_fp32_vs_fp64": {"alpha": rel(a32, a64), "dalpha_dtheta": rel(d32, d64), "adjoint": rel(g32, g64),
                             "norm": "max |x32 - x64| / max |x64| over the batch"},
        "err_fp32_vs_fp64_per_shape": {
            "alpha": max(rel(a32[b], a64[b]) for b in range(B)),
            "dalpha_dtheta": max(max(rel(d32[b, 0], d64[b, 0]), rel(d32[b, 1], d64[b, 1])) for b in range(B)),
            "norm": "max over shapes of max |x32 - x64| / max |x64| within the shape"},
        # stage 1: W's nonzero weights times the cos / sin omega k columns
        "stage1_weights_nonzero": nnz_w,
        "fwd_stage1_flops": 2.0 * nnz_w * 2 * R * B,
        "fwd_tflops_stage1_over_fwd_time": round(2.0 * nnz_w * 2 * R * B / (k_fwd * 1e-3) / 1e12, 2),
        "fp32_mfma_peak_tflops": PEAK_FP32_TFLOPS,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_projection(ks, theta[0])
        out["cpu_baseline"]["speedup_gpu_fwd_tangent"] = round(out["shapes_per_s_fwd_tangent"] /
                                                               out["cpu_baseline"]["value"], 1)
    return out


def bench_cholesky_spd(ctx, modalcmb, G, reps=3):
    """mcmb_orthonormalise and mcmb_cholesky_solve (ortho_prim_mpi.c:198-284,
    eigenmode_mpi.c:195-219) on G + 1e-4 diag(G): best of reps, host-to-host
    and the device time of the factorisation and the inverse."""
    n = G.shape[0]
    Gr = G + 1e-4 * np.diag(np.diag(G))
    rng = np.random.default_rng(7)
    alpha = rng.standard_normal((n, 4))
    best = None
    for _ in range(reps):
        t = time.time()
        try:
            ortho, lam, _ = ctx.orthonormalise(Gr)
        except modalcmb.NotSPDError as e:
            return {"status": "not positive definite in FP64: " + str(e)}
        t_o = time.time() - t
        k_ms = ctx.last_kernel_ms()[0]
        t = time.time()
        beta = ctx.cholesky_solve(ortho, alpha)
        t_s = time.time() - t
        if best is None or t_o + t_s < best[0] + best[2]:
            best = (t_o, k_ms, t_s)
    X = rng.standard_normal((n, 8))
    inv_err = float(np.abs(np.tril(lam) @ (np.tril(ortho) @ X) - X).max() / np.abs(X).max())
    res = float(np.abs(Gr @ beta - alpha).max() / np.abs(alpha).max())
    return {"n": n, "orthonormalise_s": round(best[0], 4), "factor_and_inverse_kernel_ms": round(best[1], 2),
            "solve_4_rhs_s": round(best[2], 4), "inverse_err": inv_err, "solve_residual": res,
            "note": "host to host: the Gram in, ortho and lambda (197 MB each) out; the solve takes ortho in"}


def traffic_program(config):
    """Calibration copies (the library's plain copy kernel, 1 GiB read and 1
    GiB written, three times) and one FP64 Gamma, nothing else on the device;
    one HIP runtime (the library's) in the process."""
    import modalcmb
    from modalcmb import workloads as W
    cfg = W.WORKLOADS[config]
    ws = W.lspace(cfg["L"], cfg["R"], cfg["modulated"])
    with modalcmb.Context(0) as ctx:
        gib = 1 << 30
        x, y = ctx.device_alloc(gib), ctx.device_alloc(gib)
        for _ in range(3):
            ctx.device_copy(y, x, gib)
        ctx.device_free(x)
        ctx.device_free(y)
        ctx.lspace_set(2, cfg["L"], ws["q"], ws["v"], ws["order"])
        t = time.time()
        ctx.gamma_tet(modalcmb.FP64)
        print(f"traffic program: 3 x 1 GiB copies, FP64 Gamma lmax={cfg['L']} modes={ws['order'].shape[0]} "
              f"in {time.time() - t:.3f} s", flush=True)


def polarised_gl(ctx, ws, L):
    """The polarised GL components TTE, TEE, EEE (calculate_orthol
    result[1..3]) at the workload's size: the E field takes the T basis with
    a modulated spectrum, and each component its first n rows of the
    lexicographic table.  Steady time per call into host memory (second
    call), the direct node sum's time (MCMB_GL_PATH=direct) and the
    difference between the two."""
    import modalcmb
    q, v, order = ws["q"], ws["v"], ws["order"]
    R, n = q.shape[1], order.shape[0]
    vE = v * (1.0 + 0.05 * np.cos(np.arange(len(v)) / 40.0))
    tte = np.array([(a, b, c) for a in range(R) for b in range(a, R) for c in range(R)], dtype=np.int32)[:n]
    tee = np.array([(a, b, c) for a in range(R) for b in range(R) for c in range(b, R)], dtype=np.int32)[:n]
    ctx.lspace_set_E(2, L, q, vE, tte, tee, order)
    out = {}
    try:
        for name, pol in (("TTE", modalcmb.POL_TTE), ("TEE", modalcmb.POL_TEE), ("EEE", modalcmb.POL_EEE)):
            ctx.gamma_gl_pol(pol)
            t = time.time()
            G = ctx.gamma_gl_pol(pol)
            t_c = time.time() - t
            rank = ctx.last_basis_rank()
            t = time.time()
            Gd = with_path_gl_direct(lambda: ctx.gamma_gl_pol(pol))
            t_d = time.time() - t
            d = np.sqrt(np.outer(np.abs(np.diag(Gd)), np.abs(np.diag(Gd))))
            out[name] = {"s": round(t_c, 4), "basis_rank": rank, "direct_node_sum_s": round(t_d, 4),
                         "err_vs_direct": float((np.abs(G - Gd) / d).max())}
    finally:
        ctx.lspace_clear_E()
    out["note"] = ("E field: the T basis with C_l modulated by 5%, all components 4960 modes; into host memory, "
                   "steady; err normalised by sqrt(|G_mm G_nn|)")
    return out


def with_path_gl_direct(fn):
    os.environ["MCMB_GL_PATH"] = "direct"
    try:
        return fn()
    finally:
        del os.environ["MCMB_GL_PATH"]


def with_path(path, fn):
    """Run fn() with the Gram contraction order MCMB_GRAM_PATH=path."""
    old = os.environ.get("MCMB_GRAM_PATH")
    os.environ["MCMB_GRAM_PATH"] = path
    try:
        return fn()
    finally:
        if old is None:
            del os.environ["MCMB_GRAM_PATH"]
        else:
            os.environ["MCMB_GRAM_PATH"] = old


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a step is about 3 ms: 5 warm-up steps bring the clocks to their steady state, and the
    # mean over 20 steps (and its per-phase HIP-event times) is stable to about 1%
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="lmax2500_r30",
                    help="l-space workload of the headline Gamma (a key of modalcmb.workloads.WORKLOADS)")
    ap.add_argument("--slices", type=int, default=32,
                    help="time the direct contractions on one of this many row slices and extrapolate")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-projection", action="store_true")
    ap.add_argument("--no-other-workloads", action="store_true",
                    help="skip the lmax 500 and lmax 2000 Gammas (a kernel trace then averages one workload)")
    ap.add_argument("--traffic-program", action="store_true",
                    help="for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes: three 1 GiB device copies "
                         "(calibration), then one FP64 Gamma of the workload; prints one line")
    args = ap.parse_args()
    if args.traffic_program:
        traffic_program(args.workload)
        return
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)

    import modalcmb
    from modalcmb import workloads as W

    cfg = W.WORKLOADS[args.workload]
    L, R = cfg["L"], cfg["R"]
    ws = W.lspace(L, R, cfg["modulated"])
    n = ws["order"].shape[0]
    NP = R * (R + 1) // 2
    rows = lspace_rows(L)
    ctx = modalcmb.Context(0)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    npts = ctx.npoints()
    workload = f"gamma_l{L}_n{n}"
    log(f"workload {args.workload}: lmax={L} modes={n} points={npts} rows={rows}")

    # headline: FP64 separable Gamma into a device buffer
    d_gamma = ctx.device_alloc(n * n * 8)
    # the first warm-up call on the setup also factorises the basis and
    # builds the host row tables (kept for later calls)
    t_first = None
    if args.warmup > 0:
        ctx.sync()
        t = time.time()
        ctx.gamma_tet_device(modalcmb.FP64, d_gamma)
        ctx.sync()
        t_first = round(time.time() - t, 4)
    for _ in range(max(0, args.warmup - 1)):
        ctx.gamma_tet_device(modalcmb.FP64, d_gamma)
    log(f"warmup done ({ctx.last_kernel_ms()[0]:.2f} ms kernel)")
    ctx.sync()
    kms, phases = [], {}
    t0 = time.time()
    for _ in range(args.steps):
        ctx.gamma_tet_device(modalcmb.FP64, d_gamma)
        kms.append(ctx.last_kernel_ms()[0])
        for name, ms in ctx.last_phases():
            phases.setdefault(name, []).append(ms)
    ctx.sync()
    t_step = (time.time() - t0) / args.steps
    k_step = float(np.mean(kms))
    phases = {k: round(float(np.mean(v)), 3) for k, v in phases.items()}
    G = ctx.device_to_host(d_gamma, np.empty((n, n)))
    ctx.device_free(d_gamma)
    log(f"separable FP64 Gamma: {t_step * 1e3:.2f} ms per step, {k_step:.2f} ms kernel, phases {phases}")
    rank = ctx.last_basis_rank()
    t = time.time()
    G_host = ctx.gamma_tet(modalcmb.FP64)
    t_host = time.time() - t
    assert np.array_equal(G, G_host)
    # the same separable contraction over the pair products themselves
    os.environ["MCMB_SEP_COMPRESS"] = "0"
    try:
        d_gp = ctx.device_alloc(n * n * 8)
        ctx.gamma_tet_device(modalcmb.FP64, d_gp)
        ctx.sync()
        t = time.time()
        for _ in range(2):
            ctx.gamma_tet_device(modalcmb.FP64, d_gp)
        ctx.sync()
        t_pairs = (time.time() - t) / 2
        G_pairs = ctx.device_to_host(d_gp, np.empty((n, n)))
        ctx.device_free(d_gp)
    finally:
        del os.environ["MCMB_SEP_COMPRESS"]
    log(f"separable over the pairs: {t_pairs * 1e3:.2f} ms per call")

    # the same Gram by the other contraction orders
    t = time.time()
    G_gl = ctx.gamma_gl()
    t_gl_first = time.time() - t
    t = time.time()
    G_gl = ctx.gamma_gl()
    t_gl = time.time() - t
    gl_pol = polarised_gl(ctx, ws, L)
    log("polarised GL: " + ", ".join(f"{k} {v['s'] * 1e3:.1f} ms" for k, v in gl_pol.items() if k != "note"))
    t = time.time()
    G_fac = with_path("factored", lambda: ctx.gamma_tet(modalcmb.FP64))
    t_fac = time.time() - t
    log(f"GL {t_gl:.3f} s, factored {t_fac:.2f} s")
    s_cnt = args.slices
    r_sl = s_cnt // 2
    # slices by the factored order (the separable launches stay full-size, so
    # their profiled averages are per full Gamma)
    S_fac, frac = with_path("factored", lambda: ctx.gamma_tet_slice(modalcmb.FP64, r_sl, s_cnt))
    t = time.time()
    S_dir, _ = with_path("direct", lambda: ctx.gamma_tet_slice(modalcmb.FP64, r_sl, s_cnt))
    t_dir = time.time() - t
    t = time.time()
    S_mix, _ = ctx.gamma_tet_slice(modalcmb.MIXED, r_sl, s_cnt)
    t_mix = time.time() - t
    log(f"direct slice {r_sl}/{s_cnt} ({frac:.4f}): FP64 {t_dir:.2f} s, mixed {t_mix:.2f} s")
    gamma_err = {
        "vs_fp64_separable_pairs_full": rel_elem(G, G_pairs),
        "vs_fp64_factored_full": rel_elem(G, G_fac),
        "fp64_factored_vs_direct_slice": rel_elem(S_fac, S_dir),
        "mixed_direct_vs_fp64_factored_slice": rel_elem(S_mix, S_fac),
        "vs_fp64_gauss_legendre_full": rel_elem(G, G_gl),
        "gl_note": "the GL factorised Gram carries ~1e-10 rounding of its own at lmax 2500 (it moves by that "
                   "much between 3L/2+1, 2L+1 and 3L nodes); the two tetrapyd orders agree to ~1e-13",
        "norm": "max_mn |G - G_x| / sqrt(|G_x,mm G_x,nn|)",
    }

    # late-side orthonormalisation (eigen-decomposition on the GPU, PCA branch
    # when the normalised Gram has a negative eigenvalue)
    # (the first call in a process also allocates the eigensolver's work
    # buffers, which the context keeps; the next calls show the steady cost)
    t = time.time()
    r_sep, lam_sep, _, _, e_sep = ctx.orthonormalise_late(G, pca_flag=1)
    t_late_first = time.time() - t
    t = time.time()
    r_fac, lam_fac, _, _, e_fac = ctx.orthonormalise_late(G_fac, pca_flag=1)
    t_late = time.time() - t
    r_gl, lam_gl, _, _, e_gl = ctx.orthonormalise_late(G_gl, pca_flag=1)
    log(f"late orthonormalisation (tridiagonal eigen-decomposition + PCA) {t_late:.3f} s per Gram "
        f"({t_late_first:.3f} s for the first in the process)")
    lmax = float(e_fac[0])
    cvec = np.random.default_rng(20251202).standard_normal(n)
    alpha = G_fac @ cvec
    b_sep, b_fac, b_gl = lam_sep @ alpha, lam_fac @ alpha, lam_gl @ alpha
    eig = {"lambda_max": lmax, "negative_eigenvalues_fp64": int((e_fac < 0).sum())}
    beta = {"method": "late-side PCA (ortho_late_mpi.c:784-805): the normalised Gram has negative eigenvalues",
            "rank_separable": int(r_sep), "rank_factored": int(r_fac), "orthonormalise_late_s": round(t_late, 3),
            "orthonormalise_late_first_call_s": round(t_late_first, 3),
            "orthonormalise_late_note": "4960 x 4960, host to host; Householder tridiagonalisation and divide and "
                                        "conquer on the GPU; the first call in a process also allocates the work "
                                        "buffers the context keeps"}
    for name, e_x, b_x in (("vs_factored", e_fac, b_fac), ("vs_gauss_legendre", e_gl, b_gl)):
        eig[name] = {"max_abs_err_over_lambda_max": float(np.abs(e_sep - e_x).max() / lmax)}
        beta[name] = {"rel_err_l2_all": float(np.linalg.norm(b_sep - b_x) / np.linalg.norm(b_x))}
        for thr in (1e-2, 1e-4, 1e-6):
            k = e_x > thr * lmax
            eig[name][f"rel_err_lambda_gt_{thr:g}_lmax"] = float((np.abs(e_sep[k] - e_x[k]) / e_x[k]).max())
            beta[name][f"rel_err_l2_lambda_gt_{thr:g}_lmax"] = float(np.linalg.norm(b_sep[k] - b_x[k]) /
                                                                   np.linalg.norm(b_x[k]))
    # the prim-side Cholesky (ortho_prim_mpi.c:198-284) on the same Gram
    t = time.time()
    try:
        ctx.orthonormalise(G)
        beta["cholesky_prim"] = "ok"
    except modalcmb.NotSPDError as e:
        beta["cholesky_prim"] = "not positive definite in FP64: " + str(e)
    beta["cholesky_prim_s"] = round(time.time() - t, 4)
    beta["kappa_lower_bound"] = lmax / max(float(np.abs(e_fac).min()), 1e-300)
    log("orthonormalisation done")
    cholesky_spd = bench_cholesky_spd(ctx, modalcmb, G)
    log(f"SPD Cholesky: {cholesky_spd}")
    # one-shot: a fresh setup (new generation: the basis factorisation and
    # the host row tables are rebuilt) and its first Gamma into host memory
    ctx.sync()
    t = time.time()
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    t_set = time.time() - t
    t = time.time()
    G_one = ctx.gamma_tet(modalcmb.FP64)
    t_one = time.time() - t
    parts = {name: round(ms, 3) for name, ms in ctx.last_setup()}
    assert np.array_equal(G_one, G)
    one_shot = {"lspace_set_s": round(t_set, 4), "first_gamma_to_host_s": round(t_one, 4),
                "total_s": round(t_set + t_one, 4), "first_gamma_parts_ms": parts,
                "note": "what one gamma_3d_host-style call costs on a fresh setup: tables to the device, basis "
                        "factorisation, host row tables, buffer allocation, the Gamma and its 197 MB copy "
                        "to pageable host memory"}
    log(f"one-shot: setup {t_set:.4f} s + first Gamma to host {t_one:.4f} s")
    ctx.close()

    l3_ms = phases.get("l3 contraction (row moments)")
    nb = rank if rank > 0 else NP
    l3_flops = 2.0 * npts * nb
    achieved = l3_flops / (l3_ms * 1e-3) / 1e12 if l3_ms else None
    fp64_direct_s = t_dir / frac
    mixed_direct_s = t_mix / frac
    out = {
        "metric": METRIC,
        "value": round(t_step, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"tetrapyd Gamma, FP64 separable contraction over the compressed pair basis, "
                               f"lmax={L}, modes={n} (r<=s<=t<{R}), exact 3j weights, output in HBM",
                   "lmax": L, "modes": n, "tetrapyd_points": int(npts), "tetrapyd_rows": rows, "pairs": NP},
        "gamma_kernel_ms": round(k_step, 3),
        "phases_ms": phases,
        "basis_rank": rank,
        "gamma_first_call_s": t_first,
        "one_shot": one_shot,
        "cholesky_spd": cholesky_spd,
        "fp64_separable_pairs_s": round(t_pairs, 5),
        "gamma_with_host_copy_s": round(t_host, 4),
        "fp64_factored_s": round(t_fac, 3),
        "fp64_direct_s": round(fp64_direct_s, 2),
        "mixed_direct_s": round(mixed_direct_s, 2),
        "direct_note": f"slice {r_sl} of {s_cnt} ({frac:.4f} of the point chunks) timed and extrapolated",
        "speedup_vs_fp64_direct": round(fp64_direct_s / t_step, 1),
        "speedup_vs_fp64_separable_pairs": round(t_pairs / t_step, 2),
        "speedup_vs_mixed_direct": round(mixed_direct_s / t_step, 1),
        "fp64_gl_s": round(t_gl, 4),
        "fp64_gl_first_call_s": round(t_gl_first, 4),
        "fp64_gl_pol": gl_pol,
        "mixed_note": "the separable FP64 contraction is faster than the mixed-precision direct contraction and "
                      "meets the 1e-10 bar, so the headline is FP64",
        "gamma_err": gamma_err,
        "eig": eig,
        "beta": beta,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2) if achieved else None, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 4) if achieved else None,
                     "traffic": hbm_traffic(workload),
                     "kernel": f"sep_rowmom_lw, l3 contraction over {nb} basis columns (FP64 MFMA 16x16x4, "
                               f"3j weights formed in the kernel)",
                     "algorithmic_flops_per_launch": l3_flops, "kernel_ms": l3_ms},
    }
    if not args.no_other_workloads:
        out["fp64_gamma_other_workloads"] = bench_other_workloads(modalcmb, W, args.workload)
        log("other workloads done")
    if not args.no_projection:
        out["projection"] = bench_projection(modalcmb, W, args)
        log("projection done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline_gl(ws, L)
        out["speedup_vs_cpu"] = round(out["cpu_baseline"]["value"] / t_step, 1)
        out["speedup_one_shot_vs_cpu"] = round(out["cpu_baseline"]["value"] / one_shot["total_s"], 1)
        log(f"cpu baseline (GL, whole Gamma) {out['cpu_baseline']['value']} s")
        out["cpu_baseline"]["direct_extrapolated"] = cpu_baseline(ws, L, args.cpu_seconds)
        log("cpu direct sample done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
