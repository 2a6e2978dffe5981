"""GPU: FP64 eigen-decomposition (Householder tridiagonalisation with the
own divide and conquer or rocSOLVER's, block and scalar Jacobi) and the late-side orthonormalisation with its PCA fallback,
against LAPACK (numpy eigh) and the oracle."""
import os

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def _sym(n, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n))
    return 0.5 * (a + a.T)


def _setenv(name, value):
    if value is None:
        os.environ.pop(name, None)
    else:
        os.environ[name] = value


@pytest.fixture(params=["tridiagonal", "tridiagonal-rocsolver", "block", "scalar"])
def eig_mode(request):
    """Householder tridiagonalisation with the own divide and conquer (the
    default) or with rocSOLVER's (MCMB_TRIDIAG=rocsolver), block Jacobi
    (MCMB_EIGEN=block) or the scalar Jacobi (MCMB_EIGEN=scalar)."""
    old = os.environ.get("MCMB_EIGEN"), os.environ.get("MCMB_TRIDIAG")
    tri = request.param.startswith("tridiagonal")
    _setenv("MCMB_EIGEN", None if tri else request.param)
    _setenv("MCMB_TRIDIAG", "rocsolver" if request.param == "tridiagonal-rocsolver" else None)
    yield request.param
    _setenv("MCMB_EIGEN", old[0])
    _setenv("MCMB_TRIDIAG", old[1])


@pytest.mark.parametrize("n", [1, 2, 3, 8, 33, 64, 257, 600])
def test_eigen_sym_against_lapack(ctx, n, eig_mode):
    A = _sym(n, n)
    ev, V, sweeps = ctx.eigen_sym(A)
    ref, Wl = np.linalg.eigh(A)
    ref, Wl = ref[::-1], Wl[:, ::-1]
    nrm = max(np.abs(ref).max(), 1.0)
    assert np.all(np.diff(ev) <= 0)
    assert np.abs(ev - ref).max() < 1e-13 * nrm * max(n, 4)
    assert np.abs(A @ V - V * ev).max() < 1e-12 * nrm * max(1.0, np.sqrt(n))
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12 * max(1.0, np.sqrt(n))
    imax = np.argmax(np.abs(V), axis=0)
    assert np.all(V[imax, np.arange(n)] > 0)
    gaps = np.abs(np.diff(ref))
    for k in range(n):
        sep = min(gaps[k - 1] if k > 0 else np.inf, gaps[k] if k < n - 1 else np.inf)
        if sep > 1e-4 * nrm:
            assert abs(abs(V[:, k] @ Wl[:, k]) - 1.0) < 1e-9
    if eig_mode.startswith("tridiagonal"):
        assert sweeps == 0
    else:
        assert 0 < sweeps < 20 or n == 1


@pytest.mark.parametrize("n", [97, 480])
def test_eigen_sym_graded_singular_spectrum(ctx, n, eig_mode):
    # a Gram-like spectrum: 12 decades, then entries at the FP64 noise floor
    rng = np.random.default_rng(n)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.r_[np.logspace(0, -12, n - 20), 1e-17 * rng.standard_normal(20)]
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    ev, V, sweeps = ctx.eigen_sym(A)
    ref = np.linalg.eigvalsh(A)[::-1]
    assert np.abs(ev - ref).max() < 1e-13 * n
    assert np.abs(A @ V - V * ev).max() < 1e-13 * np.sqrt(n)
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12 * np.sqrt(n)
    big = int((lam > 1e-3).sum())
    # the large eigenvalues to normwise accuracy, relative to themselves
    assert np.abs(ev[:big] / ref[:big] - 1).max() < 1e-10
    assert sweeps < 60


@pytest.mark.parametrize("n", [4000, 4960])
def test_eigen_sym_large_against_lapack(ctx, n):
    """Householder tridiagonalisation at the sizes of the lmax-2000 and
    lmax-2500 Grams: eigenvalues against LAPACK to normwise accuracy,
    residual and orthogonality of the vectors (checked on 64 random
    combinations of them, so the check itself stays O(n^2))."""
    A = _sym(n, n)
    ev, V, sweeps = ctx.eigen_sym(A)
    ref = np.linalg.eigvalsh(A)[::-1]
    nrm = np.abs(ref).max()
    assert sweeps == 0
    assert np.all(np.diff(ev) <= 0)
    assert np.abs(ev - ref).max() < 1e-13 * nrm * 8
    X = np.random.default_rng(1).standard_normal((n, 64))
    Y = V @ X
    assert np.abs(A @ Y - V @ (ev[:, None] * X)).max() < 1e-12 * nrm * np.sqrt(n) * 8
    assert np.abs(V.T @ Y - X).max() < 1e-12 * np.sqrt(n) * 8


def _wilkinson(m):
    """W_{2m+1}^+: diagonal |m - i|, unit off-diagonal (pairs of eigenvalues
    that agree to many digits)."""
    n = 2 * m + 1
    return np.diag(np.abs(np.arange(n) - m).astype(float)) + np.eye(n, k=1) + np.eye(n, k=-1)


def _hard_spectra():
    rng = np.random.default_rng(5)
    out = {"wilkinson101": _wilkinson(50)}
    # twelve W_21^+ glued by 1e-10 off-diagonals: clusters of 12 eigenvalues
    # within ~1e-10 of each other
    W21 = _wilkinson(10)
    G = np.zeros((252, 252))
    for b in range(12):
        G[21 * b:21 * b + 21, 21 * b:21 * b + 21] = W21
        if b:
            G[21 * b, 21 * b - 1] = G[21 * b - 1, 21 * b] = 1e-10
    out["glued_wilkinson252"] = G
    out["identity300"] = np.eye(300)
    # two tight clusters around 1 and 2 in a random basis
    Q, _ = np.linalg.qr(rng.standard_normal((400, 400)))
    lam = np.r_[np.ones(200), 2 * np.ones(200)] + 1e-13 * rng.standard_normal(400)
    out["clusters400"] = 0.5 * ((Q * lam) @ Q.T + ((Q * lam) @ Q.T).T)
    return out


@pytest.mark.parametrize("name", ["wilkinson101", "glued_wilkinson252", "identity300", "clusters400"])
def test_eigen_sym_tridiagonal_hard_spectra(ctx, name):
    """Spectra that exercise the divide and conquer's deflation: eigenvalue
    pairs closer than the tolerance (a Givens rotation deflates one of
    each), whole clusters and a multiple eigenvalue (the rank-one
    update's weights vanish).  Vectors orthonormal and accurate to
    working precision whatever the gaps."""
    A = _hard_spectra()[name]
    n = A.shape[0]
    for solver in ("dc", "rocsolver"):
        old = os.environ.get("MCMB_TRIDIAG")
        os.environ["MCMB_TRIDIAG"] = solver
        try:
            ev, V, sweeps = ctx.eigen_sym(A)
        finally:
            _setenv("MCMB_TRIDIAG", old)
        ref = np.linalg.eigvalsh(A)[::-1]
        nrm = np.abs(ref).max()
        assert sweeps == 0
        assert np.all(np.diff(ev) <= 0)
        assert np.abs(ev - ref).max() < 1e-14 * nrm * 16, solver
        assert np.abs(A @ V - V * ev).max() < 1e-14 * nrm * 16 * np.sqrt(n), solver
        assert np.abs(V.T @ V - np.eye(n)).max() < 1e-14 * 16 * np.sqrt(n), solver


def test_eigen_sym_repeated_calls_bitwise_equal(ctx):
    """The work buffers stay with the context between calls, and a call with
    another order reallocates them; every call gives the same bits (the
    reductions add their partial sums in a fixed order)."""
    A, B = _sym(600, 11), _sym(257, 12)
    ref = ctx.eigen_sym(A)
    runs = [ctx.eigen_sym(A), ctx.eigen_sym(B), ctx.eigen_sym(B), ctx.eigen_sym(A)]
    for r in (runs[0], runs[3]):
        assert np.array_equal(r[0], ref[0]) and np.array_equal(r[1], ref[1])
    assert np.array_equal(runs[1][0], runs[2][0]) and np.array_equal(runs[1][1], runs[2][1])
    ev, V = runs[1][0], runs[1][1]
    assert np.abs(B @ V - V * ev).max() < 1e-12 * np.abs(ev).max() * np.sqrt(257)


def test_eigen_sym_matches_oracle_vectors(ctx):
    """Same sign convention as the oracle: the eigenvectors agree entrywise."""
    A = _sym(120, 7)
    ev, V, _ = ctx.eigen_sym(A)
    st, e2, V2 = O.eigen_sym(A)
    assert st >= 0
    assert np.abs(ev - e2).max() < 1e-12
    assert np.abs(V - V2).max() < 1e-9


def test_orthonormalise_late_lmax500_pca(ctx):
    """lmax=500, 220 modes, PCA branch forced (pca_flag = 0, the reference's
    gflag_pca == 0): lambda orthonormalises Gamma on the kept directions,
    lambdainv reconstructs Gamma, and both agree with the oracle's
    restatement.  The normalised Gram has kappa ~ 1e17, so the count r of
    positive eigenvalues is itself at the FP64 noise floor: only directions
    with e > 1e-10 e_max are compared."""
    c = W.WORKLOADS["lmax500_r10"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_gl()
    r, lam, linv, norm, ev = ctx.orthonormalise_late(G, pca_flag=0)
    r2, lam2, linv2, norm2, ev2 = O.ortho_late(G, pca_flag=0)
    assert 0 < r <= 220 and 0 < r2 <= 220
    assert np.array_equal(norm, norm2)
    assert np.abs(ev - ev2).max() < 1e-13 * ev[0]
    k = ev > 1e-10 * ev[0]
    nk = int(k.sum())
    assert nk <= min(r, r2)
    P = lam[:nk] @ G @ lam[:nk].T
    assert np.abs(P - np.eye(nk)).max() < 1e-6
    assert np.abs(lam[:nk] - lam2[:nk]).max() < 1e-6 * np.abs(lam2[:nk]).max()
    Gr = linv @ linv.T
    d = np.sqrt(np.outer(np.diag(G), np.diag(G)))
    assert (np.abs(Gr - G) / d).max() < 1e-10
    assert np.all(lam[r:] == 0) and np.all(linv[:, r:] == 0)


def test_orthonormalise_late_lmax500_branch(ctx):
    """pca_flag = 1: PCA when the smallest eigenvalue is negative, else the
    Cholesky branch; on this singular-to-FP64 Gram the Cholesky may fail (the
    reference aborts in GSL there), and then the PCA branch is taken."""
    c = W.WORKLOADS["lmax500_r10"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_gl()
    ev, _, _ = ctx.eigen_sym(G / np.sqrt(np.outer(np.diag(G), np.diag(G))), vectors=False)
    r, lam, linv, norm, ev2 = ctx.orthonormalise_late(G, pca_flag=1)
    spd = True
    try:
        ctx.orthonormalise(G)
    except modalcmb.NotSPDError:
        spd = False
    if ev[-1] < 0 or not spd:
        # PCA: the same outputs as with the branch forced
        r0, lam0, linv0, _, _ = ctx.orthonormalise_late(G, pca_flag=0)
        assert r == r0 and np.array_equal(lam, lam0) and np.array_equal(linv, linv0)
    else:
        assert r == 220 and np.allclose(np.triu(lam, 1), 0)


def test_orthonormalise_late_cholesky_branch_matches_prim(ctx):
    """SPD Gram and pca_flag = 1: the Cholesky branch, equal to the lower
    triangles of mcmb_orthonormalise's symmetrised outputs."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((300, 90))
    G = X.T @ X
    r, lam, linv, norm, ev = ctx.orthonormalise_late(G, pca_flag=1)
    assert r == 90
    o, l2, n2 = ctx.orthonormalise(G)
    assert np.array_equal(linv, np.tril(o)) and np.array_equal(lam, np.tril(l2))
    assert np.allclose(lam @ linv, np.eye(90), atol=1e-11)
