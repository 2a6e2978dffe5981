"""Benchmark of the Gamma and shape-projection hot path on one MI355X.

Headline (value): seconds per mixed-precision tetrapyd Gamma at lmax=2500 with
4960 modes (r <= s <= t < 30), 653,383,126 tetrapyd points.  One step = one full
Gamma; the 1D basis, weights and mode table are resident on the device before
the timed region and the Gram comes back to the host each step.

Reported beside it:
  fp64_gpu_s          FP64 tetrapyd Gamma on the GPU (one contiguous slice of
                      the chunk list timed, extrapolated by its work fraction;
                      --fp64-full times the whole Gram)
  speedup_vs_fp64     fp64_gpu_s / value
  fp64_gl_gpu_s       the same Gamma by the Gauss-Legendre factorisation in
                      FP64 (exact 3j weights make it identical); it is the
                      exact reference for the errors below
  gamma_err_mixed     max_mn |G_mixed - G_ref| / sqrt(G_mm G_nn) and the
                      Frobenius relative error, full Gram
  eig_err_mixed       eigenvalues of C = D^-1 G D^-1 (D = sqrt diag G) from the
                      GPU Jacobi eigen-decomposition, mixed vs reference,
                      relative error over the well-determined ones
  beta                orthonormalised coefficients beta^R = lambda alpha with
                      alpha = G_ref c (seeded c), lambda from the late-side
                      orthonormalisation (PCA fallback, as the reference takes
                      when the normalised Gram has negative eigenvalues) of the
                      mixed and of the reference Gram; the prim-side Cholesky
                      status of the reference Gram
  projection          256-shape batch (omega, phi) at 300 GL nodes, 4960 modes:
                      forward with tangents (alpha, d alpha / d theta) and the
                      adjoint, FP32 MFMA with FP64 sums; gradients/s and the
                      error against the FP64 path
  roofline            the mixed Gram kernel against BF16 dense MFMA peak
                      (algorithmic flops N_tet * n * (n+1) per Gamma)
  cpu_baseline        the FP64 C oracle (same direct tetrapyd algorithm) on
                      all host cores over a bounded sample of i-rows,
                      extrapolated by point count

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--fp64-slices 16]
                       [--fp64-full] [--cpu-seconds 20] [--no-cpu] [--config 3]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modal-cmb_amd"))

METRIC = "s per Gamma (lmax 2500, 4960 modes) mixed vs FP64 GPU at 1e-10 err; grads/s"
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    d[d == 0] = 1.0
    return float((np.abs(A - B) / d).max())


def hbm_traffic(workload):
    """HBM bytes per launch of the mixed Gram kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    p = os.path.join(ROOT, "profiles", "gamma_tet_mixed_hbm_traffic.md")
    if not os.path.exists(p):
        return None
    for line in open(p):
        m = re.match(r"\s*traffic_bytes_per_gamma\[(\w+)\]\s*=\s*([0-9.eE+-]+)", line)
        if m and m.group(1) == workload:
            return float(m.group(2))
    return None


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


def bench_projection(modalcmb, W, args):
    cfg = W.CONFIGS[5]
    ks = W.kspace(cfg["N"], cfg["R"])
    theta = W.shape_batch(cfg["batch"], cfg["seed"])
    ctx = modalcmb.Context(0)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    B, n = theta.shape[0], ctx.nk
    gbar = np.random.default_rng(7).standard_normal((B, n))
    ctx.project_fwd(theta, modalcmb.FP32)     # warm-up (builds the weight cube)
    ctx.project_bwd(theta, gbar, modalcmb.FP32)
    ctx.sync()
    t = time.time()
    a32, d32 = ctx.project_fwd(theta, modalcmb.FP32)
    t_fwd = time.time() - t
    k_fwd = ctx.last_kernel_ms()[0]
    t = time.time()
    g32 = ctx.project_bwd(theta, gbar, modalcmb.FP32)
    t_bwd = time.time() - t
    k_bwd = ctx.last_kernel_ms()[0]
    a64, d64 = ctx.project_fwd(theta, modalcmb.FP64)
    g64 = ctx.project_bwd(theta, gbar, modalcmb.FP64)
    npts = ctx.kspace_npoints()
    ctx.close()

    def rel(x, y):
        return float(np.abs(x - y).max() / np.abs(y).max())

    # stage-1 contraction flops per shape: 2 (N(N+1)/2) Kp (4R): the masked
    # cube over pair rows i <= j (W is symmetric in i, j) times the four
    # angle / basis column sets per basis index; the adjoint repeats it
    N, R = cfg["N"], cfg["R"]
    Kp = (N + 15) // 16 * 16
    out = {
        "workload": f"{B} shapes S=sin(omega K + phi)/(k1k2k3)^2 form, GL {N}/side (k in [2e-4,0.2]), "
                    f"{n} modes (r<=s<=t<{R}), masked cube {npts} ordered nodes",
        "fwd_tangent_s": round(t_fwd, 4), "fwd_kernel_ms": round(k_fwd, 2),
        "adjoint_s": round(t_bwd, 4), "adjoint_kernel_ms": round(k_bwd, 2),
        "shapes_per_s_fwd_tangent": round(B / t_fwd, 1),
        "grads_per_s_fwd_tangent": round(2 * B / t_fwd, 1),
        "grads_per_s_fwd_plus_adjoint": round(2 * B / (t_fwd + t_bwd), 1),
        "err_fp32_vs_fp64": {"alpha": rel(a32, a64), "dalpha_dtheta": rel(d32, d64), "adjoint": rel(g32, g64),
                             "norm": "max |x32 - x64| / max |x64| over the batch"},
        "fwd_stage1_flops": 2.0 * (N * (N + 1) // 2) * Kp * 4 * R * B,
        "fwd_tflops_stage1_over_fwd_time": round(2.0 * (N * (N + 1) // 2) * Kp * 4 * R * B / (k_fwd * 1e-3) / 1e12, 2),
        "fp32_mfma_peak_tflops": PEAK_FP32_TFLOPS,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_projection(ks, theta[0])
        out["cpu_baseline"]["speedup_gpu_fwd_tangent"] = round(out["shapes_per_s_fwd_tangent"] /
                                                               out["cpu_baseline"]["value"], 1)
    return out


def traffic_program(config):
    """Calibration copies and one mixed Gamma, nothing else on the device.
    torch (only used for the plain copy kernel) is imported before the
    library so that the process holds one HIP runtime."""
    import torch
    x = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    del x, y
    import modalcmb
    from modalcmb import workloads as W
    cfg = W.CONFIGS[config]
    ws = W.lspace(cfg["L"], cfg["R"], cfg["modulated"])
    with modalcmb.Context(0) as ctx:
        ctx.lspace_set(2, cfg["L"], ws["q"], ws["v"], ws["order"])
        t = time.time()
        ctx.gamma_tet(modalcmb.MIXED)
        print(f"traffic program: 3 x 1 GiB copies, mixed Gamma lmax={cfg['L']} modes={ws['order'].shape[0]} "
              f"in {time.time() - t:.2f} s", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3)
    ap.add_argument("--fp64-slices", type=int, default=16,
                    help="time FP64 on one of this many contiguous slices and extrapolate")
    ap.add_argument("--fp64-full", action="store_true", help="time the full FP64 Gamma")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-projection", action="store_true")
    ap.add_argument("--traffic-program", action="store_true",
                    help="for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes: three 1 GiB device copies "
                         "(calibration), then one mixed Gamma of the config; prints one line")
    args = ap.parse_args()
    if args.traffic_program:
        traffic_program(args.config)
        return
    if args.gpus != 1:
        print(f"bench.py: this path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(multi-GPU sharding is described, not built)", file=sys.stderr)
        sys.exit(2)

    import modalcmb
    from modalcmb import workloads as W

    cfg = W.CONFIGS[args.config]
    L, R = cfg["L"], cfg["R"]
    ws = W.lspace(L, R, cfg["modulated"])
    n = ws["order"].shape[0]
    ctx = modalcmb.Context(0)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    npts = ctx.npoints()
    flops = float(npts) * n * (n + 1)
    workload = f"gamma_l{L}_n{n}"

    log(f"config {args.config}: lmax={L} modes={n} points={npts}")
    for _ in range(args.warmup):
        ctx.gamma_tet(modalcmb.MIXED)
        log(f"warmup mixed Gamma done ({ctx.last_kernel_ms()[0] / 1e3:.3f} s kernel)")
    ctx.sync()
    kms = []
    t0 = time.time()
    for _ in range(args.steps):
        g_mixed = ctx.gamma_tet(modalcmb.MIXED)
        kms.append(ctx.last_kernel_ms()[0])
        log(f"timed mixed Gamma step done ({kms[-1] / 1e3:.3f} s kernel)")
    ctx.sync()
    t_mixed = (time.time() - t0) / args.steps
    k_mixed = float(np.mean(kms)) / 1e3

    # exact FP64 reference: Gauss-Legendre factorised Gram
    t = time.time()
    g_ref = ctx.gamma_gl()
    t_gl = time.time() - t
    log(f"GL-factorised FP64 Gamma {t_gl:.3f} s")

    # FP64 tetrapyd path on the GPU
    if args.fp64_full:
        t = time.time()
        g64 = ctx.gamma_tet(modalcmb.FP64)
        fp64_s = time.time() - t
        fp64_note = "full"
        fp64_err = rel_elem(g64, g_ref)
    else:
        s = args.fp64_slices
        r = s // 2
        t = time.time()
        _, frac = ctx.gamma_tet_slice(modalcmb.FP64, r, s)
        t64 = time.time() - t
        fp64_s = t64 / frac
        fp64_note = f"extrapolated from slice {r} of {s} ({frac:.4f} of the point chunks, {t64:.2f} s)"
        fp64_err = None
    log(f"fp64 tetrapyd path: {fp64_note}")

    err = rel_elem(g_mixed, g_ref)
    frob = float(np.linalg.norm(g_mixed - g_ref) / np.linalg.norm(g_ref))

    # late-side orthonormalisation (eigen-decomposition on the GPU, PCA branch
    # when the normalised Gram has a negative eigenvalue) of both Grams
    t = time.time()
    r_ref, lam_ref, _, _, e_ref = ctx.orthonormalise_late(g_ref, pca_flag=1)
    t_late = time.time() - t
    r_mix, lam_mix, _, _, e_mix = ctx.orthonormalise_late(g_mixed, pca_flag=1)
    log(f"late orthonormalisation (Jacobi + PCA) {t_late:.2f} s per Gram")
    lmax = float(e_ref[0])
    eig = {"lambda_max": lmax, "max_abs_err_over_lambda_max": float(np.abs(e_mix - e_ref).max() / lmax),
           "negative_eigenvalues_fp64": int((e_ref < 0).sum())}
    for thr in (1e-2, 1e-4):
        k = e_ref > thr * lmax
        eig[f"rel_err_lambda_gt_{thr:g}_lmax"] = {"count": int(k.sum()),
                                                  "max": float((np.abs(e_mix[k] - e_ref[k]) / e_ref[k]).max())}
    # beta^R = lambda alpha for alpha = Gamma_ref c (seeded c): reference vs mixed
    cvec = np.random.default_rng(20251202).standard_normal(n)
    alpha = g_ref @ cvec
    b_ref, b_mix = lam_ref @ alpha, lam_mix @ alpha
    beta = {"method": "late-side PCA (ortho_late_mpi.c:784-805): the normalised Gram has negative eigenvalues",
            "rank_ref": int(r_ref), "rank_mixed": int(r_mix), "orthonormalise_late_s": round(t_late, 3),
            "rel_err_l2_all": float(np.linalg.norm(b_mix - b_ref) / np.linalg.norm(b_ref))}
    for thr in (1e-2, 1e-4, 1e-6):
        k = e_ref > thr * lmax
        beta[f"rel_err_l2_lambda_gt_{thr:g}_lmax"] = float(np.linalg.norm(b_mix[k] - b_ref[k]) /
                                                          np.linalg.norm(b_ref[k]))
    # the prim-side Cholesky (ortho_prim_mpi.c:198-284) on the same Gram
    t = time.time()
    try:
        ctx.orthonormalise(g_ref)
        beta["cholesky_prim"] = "ok"
    except modalcmb.NotSPDError as e:
        beta["cholesky_prim"] = "not positive definite in FP64: " + str(e)
    beta["cholesky_prim_s"] = round(time.time() - t, 4)
    beta["kappa_lower_bound"] = lmax / max(float(np.abs(e_ref).min()), 1e-300)
    log("orthonormalisation done")
    ctx.close()

    achieved = flops / k_mixed / 1e12
    out = {
        "metric": METRIC,
        "value": round(t_mixed, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_mixed * 1e3, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic",
        "config": {"workload": f"tetrapyd Gamma, mixed precision (bf16 hi+lo x3, fp32 accumulate, fp64 sums), "
                               f"lmax={L}, modes={n} (r<=s<=t<{R}), exact 3j weights",
                   "lmax": L, "modes": n, "tetrapyd_points": int(npts)},
        "fp64_gpu_s": round(fp64_s, 3),
        "fp64_gpu_note": fp64_note,
        "speedup_vs_fp64": round(fp64_s / t_mixed, 3),
        "fp64_gl_gpu_s": round(t_gl, 4),
        "gamma_err_mixed": {"max_elem_rel": err, "frob_rel": frob,
                            "vs": "FP64 Gauss-Legendre factorised Gram (full)"},
        "eig_err_mixed": eig,
        "beta": beta,
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4),
                     "traffic": hbm_traffic(workload),
                     "kernel": "gram_tet_mixed_kernel (diagonal + off-diagonal tile launches)",
                     "algorithmic_flops_per_gamma": flops, "kernel_s": round(k_mixed, 4)},
    }
    if fp64_err is not None:
        out["gamma_err_fp64_tet"] = fp64_err
    if not args.no_projection:
        out["projection"] = bench_projection(modalcmb, W, args)
        log("projection done")
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ws, L, args.cpu_seconds)
        log("cpu baseline done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
