"""GPU FP64 orthonormalisation and Cholesky solve against the oracle
(ortho_prim_mpi.c:198-284 layouts) and against the invariants they must meet."""
import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def spd(n, seed, cond=1e4):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = np.logspace(0, np.log10(cond), n)
    d = rng.uniform(0.5, 3.0, n)
    return (Q * ev) @ Q.T * np.outer(d, d)


@pytest.mark.parametrize("n", [1, 5, 31, 32, 33, 63, 64, 65, 100, 257, 700, 1030])
def test_orthonormalise_matches_oracle(ctx, n):
    G = spd(n, n)
    ortho, lam, norm = ctx.orthonormalise(G)
    st, o_ref, l_ref, n_ref = O.orthonormalise(G)
    assert st == 0
    assert np.allclose(norm, n_ref, rtol=1e-15, atol=0)
    assert np.allclose(ortho, o_ref, rtol=1e-12, atol=1e-12 * np.abs(o_ref).max())
    assert np.allclose(lam, l_ref, rtol=1e-9, atol=1e-11 * np.abs(l_ref).max())
    F, Lam = np.tril(ortho), np.tril(lam)
    assert np.allclose(F @ F.T, G, rtol=1e-11, atol=1e-12 * np.abs(G).max())
    assert np.abs(Lam @ G @ Lam.T - np.eye(n)).max() < 1e-9


def test_orthonormalise_not_spd_reports_pivot(ctx):
    G = spd(40, 1)
    G[17, :] = G[16, :]
    G[:, 17] = G[:, 16]
    with pytest.raises(modalcmb.NotSPDError) as e:
        ctx.orthonormalise(G)
    assert "pivot 17" in str(e.value)


@pytest.mark.parametrize("n,nrhs", [(300, 7), (300, 1), (611, 70), (129, 130)])
def test_cholesky_solve_many_rhs(ctx, n, nrhs):
    G = spd(n, 9, cond=1e3)
    ortho, _, _ = ctx.orthonormalise(G)
    rng = np.random.default_rng(2)
    C = rng.standard_normal((n, nrhs))
    A = G @ C
    B = ctx.cholesky_solve(ortho, A)
    assert np.allclose(B, C, rtol=1e-9, atol=1e-9)
    for j in range(0, nrhs, max(1, nrhs // 7)):
        assert np.allclose(B[:, j], O.cholesky_solve(ortho, A[:, j]), rtol=1e-10, atol=1e-10)


def test_orthonormalise_inverse_is_exact_triangle(ctx):
    """tril(lambda) is (D L)^-1: lower triangular with L^-1 D^-1 on it, so
    tril(lambda) @ tril(ortho) = I to rounding, on a size with ragged tiles."""
    n = 517
    G = spd(n, 5, cond=1e3)
    ortho, lam, _ = ctx.orthonormalise(G)
    P = np.tril(lam) @ np.tril(ortho)
    assert np.abs(P - np.eye(n)).max() < 1e-11


def test_lmax500_gram_orthonormalisation_status(ctx):
    # The lmax-500 Gram (r<=s<=t<10 Legendre modes, weight h^2/C^3) has
    # kappa(D^-1 G D^-1) ~ 1e17 > 1/eps: whether FP64 Cholesky meets a
    # non-positive pivot depends on summation order (the oracle passes, the GPU
    # order stops near column 174).  Either outcome must be clean: a NOT_SPD
    # status (the reference would abort inside GSL), or a factor that
    # reproduces the normalised Gram to FP64 accuracy.
    ws = W.lspace(500, 10)
    ctx.lspace_set(2, 500, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.FP64)
    d = np.sqrt(np.diag(G))
    ev = np.linalg.eigvalsh(G / np.outer(d, d))
    assert ev.max() / max(ev.min(), 1e-300) > 1e15 or ev.min() <= 0
    try:
        ortho, lam, norm = ctx.orthonormalise(G)
    except modalcmb.NotSPDError as e:
        assert "pivot" in str(e)
    else:
        F = np.tril(ortho)
        assert np.abs(F @ F.T - G).max() < 1e-12 * np.abs(G).max()
