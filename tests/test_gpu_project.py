"""GPU k-space Gram and feature-shape projection (value, tangent, adjoint)
against the FP64 oracle and central finite differences of the oracle."""
import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    return float((np.abs(A - B) / d).max())


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def oracle_count(kn):
    N = len(kn)
    return sum(1 for a in range(N) for b in range(a, N) for c in range(b, N) if kn[c] <= kn[a] + kn[b])


@pytest.mark.parametrize("N,R", [(12, 3), (30, 5), (64, 8)])
def test_kspace_points_and_gram(ctx, N, R):
    ks = W.kspace(N, R)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    assert ctx.kspace_npoints() == oracle_count(ks["kn"])
    G = ctx.kgram(modalcmb.FP64)
    ref = O.gram_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"])
    assert rel_elem(G, ref) < 1e-12
    Gm = ctx.kgram(modalcmb.MIXED)
    assert rel_elem(Gm, ref) < 1e-5


THETAS = np.array([[150.0, 0.7], [10.0, 0.0], [499.0, 6.1], [37.5, 3.3]])


# R > 16 fills both 16-row tiles of the basis index in the MFMA contractions;
# N = 33 and 20 leave partial node chunks (of 32) and an odd count of K chunks
@pytest.mark.parametrize("N,R", [(24, 4), (60, 7), (33, 20), (20, 32)])
def test_projection_fwd_vs_oracle(ctx, N, R):
    ks = W.kspace(N, R)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    a, d = ctx.project_fwd(THETAS, modalcmb.FP64)
    a32, d32 = ctx.project_fwd(THETAS, modalcmb.FP32)
    for b, (om, ph) in enumerate(THETAS):
        ra, rom, rph = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om, ph)
        assert relmax(a[b], ra) < 1e-12
        assert relmax(d[b, 0], rom) < 1e-12
        assert relmax(d[b, 1], rph) < 1e-12
        assert relmax(a32[b], ra) < 5e-6
        assert relmax(d32[b, 0], rom) < 5e-6 and relmax(d32[b, 1], rph) < 5e-6


def test_projection_tangent_vs_central_differences(ctx):
    # the single-shape protocol (omega=150, phi=0.7, h=1e-4) at a reduced grid
    ks = W.kspace(40, 6)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    om, ph, h = 150.0, 0.7, 1e-4
    _, d = ctx.project_fwd([[om, ph]], modalcmb.FP64)
    f = lambda o, p: O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], o, p)[0]
    fd_om = (f(om + h, ph) - f(om - h, ph)) / (2 * h)
    fd_ph = (f(om, ph + h) - f(om, ph - h)) / (2 * h)
    assert relmax(d[0, 0], fd_om) < 1e-6
    assert relmax(d[0, 1], fd_ph) < 1e-6


@pytest.mark.parametrize("N,R", [(50, 6), (33, 20)])
def test_projection_adjoint_equals_jacobian_contraction(ctx, N, R):
    ks = W.kspace(N, R)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    th = W.shape_batch(40, seed=5)
    rng = np.random.default_rng(0)
    gbar = rng.standard_normal((40, ks["order"].shape[0]))
    _, d = ctx.project_fwd(th, modalcmb.FP64)
    ref = np.einsum("bn,bxn->bx", gbar, d)
    dt = ctx.project_bwd(th, gbar, modalcmb.FP64)
    assert np.abs(dt - ref).max() < 1e-11 * np.abs(ref).max()
    dt32 = ctx.project_bwd(th, gbar, modalcmb.FP32)
    assert np.abs(dt32 - ref).max() < 1e-5 * np.abs(ref).max()


def test_projection_beta_solve_roundtrip(ctx):
    # beta = Gamma_k^-1 alpha through the GPU Cholesky (eigenmode_mpi.c:195-219)
    ks = W.kspace(40, 3)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    G = ctx.kgram(modalcmb.FP64)
    ortho, lam, norm = ctx.orthonormalise(G)
    a = ctx.project_fwd([[150.0, 0.7]], modalcmb.FP64, derivatives=False)[0]
    beta = ctx.cholesky_solve(ortho, a)
    st, o_ref, _, _ = O.orthonormalise(O.gram_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"]))
    ref = O.cholesky_solve(o_ref, O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], 150.0, 0.7)[0])
    assert st == 0
    assert relmax(beta, ref) < 1e-8
    assert np.allclose(G @ beta, a, rtol=1e-8, atol=1e-10 * np.abs(a).max())


@pytest.mark.parametrize("prec", [modalcmb.FP64, modalcmb.FP32])
def test_projection_adjoint_reuses_forward_products(ctx, prec):
    # a backward call right after a forward call on the same shapes reuses the
    # forward's stage-1 products; the result is bit-identical to recomputing
    # them (forward on other shapes first), and several batches of 32 shapes
    # are covered
    ks = W.kspace(40, 5)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    th = W.shape_batch(70, seed=11)
    other = W.shape_batch(70, seed=12)
    gbar = np.random.default_rng(1).standard_normal((70, ks["order"].shape[0]))
    ctx.project_fwd(th, prec)
    reused = ctx.project_bwd(th, gbar, prec)
    ctx.project_fwd(other, prec)
    recomputed = ctx.project_bwd(th, gbar, prec)
    assert np.array_equal(reused, recomputed)
    # a new setup drops the cached products
    ctx.project_fwd(th, prec)
    ks2 = W.kspace(40, 4)
    ctx.kspace_set(ks2["kn"], ks2["kw"], ks2["q"], ks2["order"])
    g2 = gbar[:, :ks2["order"].shape[0]]
    _, d2 = ctx.project_fwd(other, modalcmb.FP64)
    ref = np.einsum("bn,bxn->bx", g2, ctx.project_fwd(th, modalcmb.FP64)[1])
    ctx.project_fwd(other, prec)
    dt = ctx.project_bwd(th, g2, prec)
    assert np.abs(dt - ref).max() < (1e-11 if prec == modalcmb.FP64 else 1e-5) * np.abs(ref).max()


def test_projection_transfers_over_more_batches_than_slots(ctx, monkeypatch):
    # 300 shapes are 10 batches of 32, more than the 8 pinned slots that carry
    # the results out and gbar in while later batches compute: the values
    # equal those of plain whole-array copies bit for bit, and the last
    # batches' shapes match the oracle
    ks = W.kspace(24, 4)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    th = W.shape_batch(300, seed=21)
    gbar = np.random.default_rng(2).standard_normal((300, ks["order"].shape[0]))
    a, d = ctx.project_fwd(th, modalcmb.FP64)
    g = ctx.project_bwd(th, gbar, modalcmb.FP64)
    a_only = ctx.project_fwd(th, modalcmb.FP64, derivatives=False)
    monkeypatch.setenv("MCMB_HOST_COPY", "direct")
    a2, d2 = ctx.project_fwd(th, modalcmb.FP64)
    g2 = ctx.project_bwd(th, gbar, modalcmb.FP64)
    assert np.array_equal(a, a2) and np.array_equal(d, d2) and np.array_equal(g, g2)
    assert np.array_equal(a_only, a)
    for b in (0, 255, 256, 299):
        ra, rom, rph = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], th[b, 0], th[b, 1])
        assert relmax(a[b], ra) < 1e-12
        assert relmax(d[b, 0], rom) < 1e-12 and relmax(d[b, 1], rph) < 1e-12
