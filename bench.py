"""Benchmark of the Gamma hot path on one MI355X.

Metric: seconds per mixed-precision tetrapyd Gamma at lmax=2500, 4960 modes
(r <= s <= t < 30), against the FP64 GPU path on the same Gram, with the error
of the mixed result stated.  One step = one full Gamma over the 653,383,126
tetrapyd points.  Inputs (1D basis, weights, mode table) are resident on the
device before the timed region; the Gram comes back to the host each step.

Extra fields:
  fp64_gpu_s            FP64 GPU Gamma time (full or extrapolated from a slice)
  speedup_vs_fp64       fp64_gpu_s / value
  gamma_err_mixed       max_mn |G_mixed - G_fp64| / sqrt(G_mm G_nn) on the
                        compared slice, and its Frobenius relative error
  roofline              the mixed Gram kernel against BF16 dense MFMA peak
                        (algorithmic flops N_tet * n * (n+1) per Gamma)
  cpu_baseline          the FP64 C oracle (same direct tetrapyd algorithm) on
                        all host cores over a bounded sample of i-rows,
                        extrapolated by point count

Usage: python bench.py [--gpus 1] [--steps 1] [--warmup 1] [--fp64-slices 16]
                       [--fp64-full] [--cpu-seconds 20] [--no-cpu] [--config 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "modal-cmb_amd"))

METRIC = "s per Gamma (lmax 2500, 4960 modes) mixed vs FP64 GPU at 1e-10 err; grads/s"
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP64_TFLOPS = 78.6


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    d[d == 0] = 1.0
    return float((np.abs(A - B) / d).max())


def cpu_baseline(ws, L, target_s):
    """Time the oracle's direct tetrapyd Gram over the top i-rows of T(2, L)
    (about target_s seconds of work on all host cores) and scale by points."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    n = ws["order"].shape[0]
    # points per i-row of T(2, L)
    cnt = np.zeros(L + 2, dtype=np.int64)
    for i in range(2, L + 1):
        j = np.arange(i, L - (i & 1) + 1)
        ke = np.minimum(i + j, L)
        kb = j + (i & 1)
        cnt[i] = int(np.maximum(0, (ke - kb) // 2 + 1).sum())
    total = int(cnt.sum())
    nthreads = O.lib().or_max_threads()
    # calibrate: rows from the top until ~a short probe, then size the sample
    def run(i_lo, i_hi):
        t = time.time()
        O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"], 0, i_lo, i_hi)
        return time.time() - t
    i_lo = L + 1 - 2 * max(16, nthreads)
    pts = int(cnt[i_lo:].sum())
    t_probe = run(i_lo, L + 1)
    rate = pts / t_probe
    want = rate * target_s
    i_lo2 = L + 1
    acc = 0
    while i_lo2 > 2 and acc + cnt[i_lo2 - 1] <= want:
        i_lo2 -= 1
        acc += int(cnt[i_lo2])
    i_lo2 = min(i_lo2, i_lo)
    pts2 = int(cnt[i_lo2:].sum())
    t2 = run(i_lo2, L + 1)
    return {
        "value": t2 * total / pts2,
        "unit": "s per Gamma (extrapolated)",
        "cores": int(nthreads),
        "kind": "port",
        "sample": f"oracle direct tetrapyd FP64 Gram, i-rows {i_lo2}..{L} of T(2,{L}): {pts2} of {total} points "
                  f"({t2:.1f} s), scaled by point count; n={n}",
    }


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
    args = ap.parse_args()
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

    # FP64 GPU path
    if args.fp64_full:
        t = time.time()
        g64 = ctx.gamma_tet(modalcmb.FP64)
        fp64_s = time.time() - t
        fp64_note = "full"
        err = rel_elem(g_mixed, g64)
        frob = float(np.linalg.norm(g_mixed - g64) / np.linalg.norm(g64))
        cmp_note = "full Gram"
    else:
        s = args.fp64_slices
        r = s // 2
        t = time.time()
        g64s, frac = ctx.gamma_tet_slice(modalcmb.FP64, r, s)
        t64 = time.time() - t
        fp64_s = t64 / frac
        fp64_note = f"extrapolated from slice {r} of {s} ({frac:.4f} of the point chunks, {t64:.2f} s)"
        gms, _ = ctx.gamma_tet_slice(modalcmb.MIXED, r, s)
        err = rel_elem(gms, g64s)
        frob = float(np.linalg.norm(gms - g64s) / np.linalg.norm(g64s))
        cmp_note = f"slice {r} of {s}"
    log(f"fp64 path: {fp64_note}")

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
        "gamma_err_mixed": {"max_elem_rel": err, "frob_rel": frob, "on": cmp_note},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_BF16_TFLOPS, 4), "traffic": None,
                     "kernel": "gram_tet_mixed_kernel (diag + off-diag launches)",
                     "algorithmic_flops_per_gamma": flops, "kernel_s": round(k_mixed, 4)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ws, L, args.cpu_seconds)
        log("cpu baseline done")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
