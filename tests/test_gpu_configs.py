"""Parity at the benchmark workloads' full sizes, through the C ABI, against
the FP64 oracle.

* lmax 2000 / 1540 modes: the first and last 10^6 points of the device
  enumeration against the walk of get_ijk_next_XXX; the whole FP64 Gram
  against the oracle's Gauss-Legendre Gram (an independent algorithm); the
  i-rows l1 >= 1900 against the oracle's direct tetrapyd sum; the FP64
  Cholesky outcome on the Gram and on a regularised Gram.
* lmax 2500 / 4960 modes: the i-rows l1 >= 2430 against the oracle's direct
  sum; the whole separable Gram against the GPU Gauss-Legendre Gram.
* 200-node GL k-grid, basis width 20 (1540 modes), omega = 150, phi = 0.7:
  alpha and d alpha / d (omega, phi) in FP64 and FP32 against the oracle, and
  the FP64 tangent against central differences (h = 1e-4) of the oracle.
* 300-node GL k-grid, basis width 30 (4960 modes), 256 seeded shapes: 4 of
  them against the oracle in FP32 and FP64, the whole batch FP32 against
  FP64, and the adjoint against the Jacobian contraction.

Tolerances are stated per assertion."""
import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def rel_elem(A, B):
    """max_mn |A - B| / sqrt(B_mm B_nn)"""
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    return float((np.abs(A - B) / d).max())


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def lspace(ctx, name):
    c = W.WORKLOADS[name]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    return c, ws


# ---- lmax 2000, 1540 modes ------------------------------------------------

def test_lmax2000_enumeration_head_and_tail_bit_exact(ctx):
    c, _ = lspace(ctx, "lmax2000_r20")
    N = ctx.npoints()
    assert N == 334831501                               # process.c:65
    head = ctx.points_at(0, 1_000_000)
    assert np.array_equal(head, O.walk_XXX_window(2, 2000, 0, 1_000_000))
    tail = ctx.points_at(N - 1_000_000, 1_000_000)
    assert np.array_equal(tail, O.walk_XXX_window(2, 2000, N - 1_000_000, 1_000_000))
    assert tuple(tail[-1]) == (2000, 2000, 2000)
    mid = N // 2
    assert np.array_equal(ctx.points_at(mid, 4096), O.walk_XXX_window(2, 2000, mid, 4096))


def test_lmax2000_gram_vs_oracle_gl_full(ctx):
    c, ws = lspace(ctx, "lmax2000_r20")
    G = ctx.gamma_tet(modalcmb.FP64)
    assert G.shape == (1540, 1540) and np.array_equal(G, G.T)
    ref = O.gamma_gl(2, c["L"], ws["q"], ws["v"], ws["order"])
    # the FP64 bar of the workload: 1e-10 relative per element
    assert rel_elem(G, ref) < 1e-10


def test_lmax2000_top_rows_vs_oracle_direct(ctx):
    c, ws = lspace(ctx, "lmax2000_r20")
    G = ctx.gamma_tet_lrange(modalcmb.FP64, 1900, 2001)
    ref = O.gamma_tet(2, c["L"], ws["q"], ws["v"], ws["order"], i_lo=1900, i_hi=2001)
    assert rel_elem(G, ref) < 1e-12
    # ranges that tile [2, lmax] sum to the whole Gram
    rest = ctx.gamma_tet_lrange(modalcmb.FP64, 0, 1900)
    assert rel_elem(G + rest, ctx.gamma_tet(modalcmb.FP64)) < 1e-13


def test_lmax2000_cholesky_outcome(ctx):
    c, ws = lspace(ctx, "lmax2000_r20")
    G = ctx.gamma_tet(modalcmb.FP64)
    # singular to FP64: the prim-side Cholesky (ortho_prim_mpi.c:232) fails on
    # both sides, the GPU with a clean status instead of GSL's abort
    with pytest.raises(modalcmb.NotSPDError):
        ctx.orthonormalise(G)
    assert O.orthonormalise(G)[0] == -3
    # regularised (relative shift 1e-4 of the diagonal): an SPD Gram of the
    # same size, condition ~1e6, factored on both sides
    Gr = G + 1e-4 * np.diag(np.diag(G))
    ortho, lam, norm = ctx.orthonormalise(Gr)
    st, o_ref, l_ref, n_ref = O.orthonormalise(Gr)
    assert st == 0
    assert np.array_equal(norm, n_ref) or relmax(norm, n_ref) < 1e-15
    assert relmax(ortho, o_ref) < 1e-10
    assert relmax(lam, l_ref) < 1e-8
    # beta = Gamma^-1 alpha for alpha = Gamma c
    cvec = np.random.default_rng(3).standard_normal(1540)
    beta = ctx.cholesky_solve(ortho, Gr @ cvec)
    assert relmax(beta, cvec) < 1e-8
    assert relmax(beta, O.cholesky_solve(o_ref, Gr @ cvec)) < 1e-9


# ---- lmax 2500, 4960 modes ------------------------------------------------

def test_lmax2500_top_rows_vs_oracle_direct(ctx):
    c, ws = lspace(ctx, "lmax2500_r30")
    G = ctx.gamma_tet_lrange(modalcmb.FP64, 2430, 2501)
    ref = O.gamma_tet(2, c["L"], ws["q"], ws["v"], ws["order"], i_lo=2430, i_hi=2501)
    assert G.shape == (4960, 4960)
    assert rel_elem(G, ref) < 1e-12


def test_lmax2500_separable_vs_gl(ctx):
    c, ws = lspace(ctx, "lmax2500_r30")
    G = ctx.gamma_tet(modalcmb.FP64)
    assert ctx.last_basis_rank() == 59
    Ggl = ctx.gamma_gl()
    # the GL Gram carries its own ~2e-10 rounding at this size
    assert rel_elem(G, Ggl) < 5e-10
    assert np.array_equal(G, G.T)


def test_lmax2500_host_copy_staged_equals_direct(ctx):
    # 197 MB outputs reach the host through pinned 16 MiB halves and host
    # threads; the bytes are those of a plain hipMemcpy, also for a length
    # that is not a multiple of the halves
    import os
    c, ws = lspace(ctx, "lmax2500_r30")
    G = ctx.gamma_tet(modalcmb.FP64)
    n = G.shape[0]
    dev = ctx.device_alloc(8 * n * n)
    try:
        ctx.gamma_tet_device(modalcmb.FP64, dev)
        os.environ["MCMB_HOST_COPY"] = "direct"
        try:
            G2 = ctx.gamma_tet(modalcmb.FP64)
            part_direct = ctx.device_to_host(dev, np.empty(n * n - 7))
        finally:
            del os.environ["MCMB_HOST_COPY"]
        part = ctx.device_to_host(dev, np.empty(n * n - 7))
    finally:
        ctx.device_free(dev)
    assert np.array_equal(G, G2)
    assert np.array_equal(part, part_direct) and np.array_equal(part, G.ravel()[:n * n - 7])


# ---- 200-node GL k-grid, basis width 20 -----------------------------------

def test_gl200_projection_vs_oracle_and_central_differences(ctx):
    c = W.WORKLOADS["gl200_r20"]
    ks = W.kspace(c["N"], c["R"])
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    om, ph, h = c["omega"], c["phi"], 1e-4
    th = np.array([[om, ph], [om + h, ph], [om - h, ph], [om, ph + h], [om, ph - h]])
    ra, rom, rph = O.project_kgl_batch(ks["kn"], ks["kw"], ks["q"], ks["order"], th)
    a, d = ctx.project_fwd(th[:1], modalcmb.FP64)
    a32, d32 = ctx.project_fwd(th[:1], modalcmb.FP32)
    assert relmax(a[0], ra[0]) < 1e-12
    assert relmax(d[0, 0], rom[0]) < 1e-12 and relmax(d[0, 1], rph[0]) < 1e-12
    # FP32 k-sum, FP64 mode sums
    assert relmax(a32[0], ra[0]) < 1e-6
    assert relmax(d32[0, 0], rom[0]) < 1e-6 and relmax(d32[0, 1], rph[0]) < 1e-6
    # central differences of the oracle (truncation ~ h^2 K^3 / 6)
    fd_om = (ra[1] - ra[2]) / (2 * h)
    fd_ph = (ra[3] - ra[4]) / (2 * h)
    assert relmax(d[0, 0], fd_om) < 1e-7 and relmax(d[0, 1], fd_ph) < 1e-7
    assert relmax(d32[0, 0], fd_om) < 2e-6 and relmax(d32[0, 1], fd_ph) < 2e-6
    # the Gram of this basis is singular to FP64 (smallest eigenvalues of the
    # normalised Gram ~ -1e-14): the Cholesky beta solve reports it on both sides
    Gk = ctx.kgram(modalcmb.FP64)
    with pytest.raises(modalcmb.NotSPDError):
        ctx.orthonormalise(Gk)
    assert O.orthonormalise(Gk)[0] == -3


# ---- 300-node GL k-grid, basis width 30, 256 shapes -----------------------

def test_gl300_batch256_vs_oracle(ctx):
    c = W.WORKLOADS["gl300_r30_batch256"]
    ks = W.kspace(c["N"], c["R"])
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    th = W.shape_batch(c["batch"], seed=c["seed"])
    a64, d64 = ctx.project_fwd(th, modalcmb.FP64)
    a32, d32 = ctx.project_fwd(th, modalcmb.FP32)
    pick = [0, 77, 150, 255]
    ra, rom, rph = O.project_kgl_batch(ks["kn"], ks["kw"], ks["q"], ks["order"], th[pick])
    for i, b in enumerate(pick):
        assert relmax(a64[b], ra[i]) < 1e-12
        assert relmax(d64[b, 0], rom[i]) < 1e-12 and relmax(d64[b, 1], rph[i]) < 1e-12
        # FP32 k-sum over 300 nodes: up to ~1e-6 of a shape's largest
        # coefficient (oscillatory shapes cancel); FP64 is the path within 1e-8
        assert relmax(a32[b], ra[i]) < 2e-6
        assert relmax(d32[b, 0], rom[i]) < 2e-6 and relmax(d32[b, 1], rph[i]) < 2e-6
    # whole batch, FP32 against FP64: per shape as above, and within 1e-8 of
    # the batch's largest coefficient
    for b in range(th.shape[0]):
        assert relmax(a32[b], a64[b]) < 2e-6
        assert relmax(d32[b, 0], d64[b, 0]) < 2e-6 and relmax(d32[b, 1], d64[b, 1]) < 2e-6
    assert relmax(a32, a64) < 1e-8 and relmax(d32, d64) < 1e-8
    # adjoint = contraction of the tangent Jacobian
    gbar = np.random.default_rng(9).standard_normal(a64.shape)
    ref = np.einsum("bn,bxn->bx", gbar, d64)
    dt64 = ctx.project_bwd(th, gbar, modalcmb.FP64)
    dt32 = ctx.project_bwd(th, gbar, modalcmb.FP32)
    assert np.abs(dt64 - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(dt32 - ref).max() < 1e-8 * np.abs(ref).max()
