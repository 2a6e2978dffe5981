"""Pin the FP64 CPU oracle (oracle/) against golden fixtures, known answers and
mathematics that a wrong restatement would violate.  CPU only."""
import json
import os

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O
from conftest import GOLDEN
from modalcmb import workloads as W


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_count_known_answer():
    # process.c:47-48,65 : record count of the l1<=l2<=l3<=2000 tetrapyd from l=2
    assert O.lib().or_count_tetrapyd_XXX(2, 2000) == 334831501


@pytest.mark.parametrize("rec", load("walk_counts.json"))
def test_walk_matches_brute_force(rec):
    lmin, lmax = rec["lmin"], rec["lmax"]
    pts = O.walk_XXX(lmin, lmax)
    assert len(pts) == rec["count"]
    brute = [(i, j, k) for i in range(lmin, lmax + 1) for j in range(i, lmax + 1)
             for k in range(j, lmax + 1) if k <= i + j and (i + j + k) % 2 == 0]
    assert [tuple(p) for p in pts] == brute  # walk order = lexicographic (i, j, k)


def test_walk_prim_matches_brute_force():
    for mx in (5, 8, 13):
        pts = O.walk_prim(0, mx)
        brute = [(i, j, k) for i in range(0, mx + 1) for j in range(i, mx + 1)
                 for k in range(j, mx + 1) if k <= i + j + 2 and (i < mx or (j == mx and k == mx))]
        assert [tuple(p) for p in pts] == brute


def test_decompose_covers_walk():
    pts = O.walk_XXX(2, 60)
    for nranks, nthreads in [(1, 1), (1, 7), (3, 4), (5, 2)]:
        seen = []
        for r in range(nranks):
            blk = O.decompose_XXX(r, nranks, nthreads, 2, 60)
            for b in blk:
                start = int(np.flatnonzero((pts == b[:3]).all(1))[0])
                seen.extend(range(start, start + b[3]))
        assert seen == list(range(len(pts)))


@pytest.mark.parametrize("rec", load("threej.json"))
def test_threej_and_h2_against_mpmath(rec):
    l1, l2, l3 = rec["l"]
    a = O.lib().or_threej000_sq(l1, l2, l3)
    b = float(mp.mpf(rec["threej_sq"]))
    assert a == pytest.approx(b, rel=2e-15, abs=0)
    assert O.lib().or_h2_exact(l1, l2, l3) == pytest.approx(float(mp.mpf(rec["h2"])), rel=2e-15)


def test_threej_sum_rule():
    # sum_{l3} (2 l3 + 1) (l1 l2 l3; 0 0 0)^2 = 1 over the allowed l3
    for l1, l2 in [(2, 2), (7, 30), (100, 137), (1200, 1300), (2500, 2499)]:
        s = sum((2 * l3 + 1) * O.lib().or_threej000_sq(l1, l2, l3)
                for l3 in range(abs(l1 - l2), l1 + l2 + 1))
        assert s == pytest.approx(1.0, abs=1e-13)


def test_geometric_asymptotic_close_to_exact_at_large_l():
    # calculate_geometric.c is asymptotic: ~1e-2 at l ~ 2, < 1e-6 at l ~ 500
    assert abs(O.lib().or_geometric_asym(500, 500, 500) / O.lib().or_h2_exact(500, 500, 500) - 1) < 1e-6
    assert abs(O.lib().or_geometric_asym(2, 2, 2) / O.lib().or_h2_exact(2, 2, 2) - 1) > 1e-3
    assert O.lib().or_geometric_asym(2, 3, 4) == 0.0 and O.lib().or_geometric_asym(2, 3, 7) == 0.0


def test_permsix():
    p = O.lib().or_permsix
    assert (p(3, 3, 3), p(2, 3, 3), p(3, 3, 2), p(2, 3, 2), p(1, 2, 3)) == (1, 3, 3, 3, 6)


@pytest.mark.parametrize("n", [5, 20, 64])
def test_gauss_legendre_nodes(n):
    ref = load("gl.json")[str(n)]
    x, w = O.gauss_legendre(n)
    xr = np.array([float(mp.mpf(s)) for s in ref["x"]])
    wr = np.array([float(mp.mpf(s)) for s in ref["w"]])
    assert np.abs(x - xr).max() < 2e-16 and np.abs(w - wr).max() < 4e-16
    # exact on monomials up to degree 2n-1
    for d in (0, 1, 2, 2 * n - 2, 2 * n - 1):
        exact = 0.0 if d % 2 else 2.0 / (d + 1)
        assert np.dot(w, x ** d) == pytest.approx(exact, abs=1e-14)


def test_gauss_legendre_large_n():
    mp.mp.dps = 40
    for n in (751, 3751):
        x, w = O.gauss_legendre(n)
        xn, wn = np.polynomial.legendre.leggauss(n)
        assert np.abs(x - xn).max() < 1e-14 and np.abs(w - wn).max() / wn.max() < 1e-7
        assert w.sum() == pytest.approx(2.0, abs=1e-13)
        if n > 1000:
            continue
        for idx in (0, 1, n // 3):   # end weights are where numpy's rule is weakest
            r = mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(x[idx]))
            d = mp.diff(lambda t: mp.legendre(n, t), r)
            wr = 2 / ((1 - r * r) * d * d)
            assert abs(x[idx] - float(r)) < 1e-16 and abs(w[idx] / float(wr) - 1) < 2e-14


def test_legendre_basis_against_mpmath():
    recs = load("legendre.json")
    x01 = np.array([(r["x"] + 1) / 2 for r in recs])
    tab = O.basis_legendre(x01, 29, 0.0, 1.0)
    for row, rec in zip(tab, recs):
        ref = np.array([float(mp.mpf(s)) for s in rec["P"]])
        assert np.abs(row - ref).max() < 2e-14   # three-term recurrence, r < 30
    # the workload generator is bit-identical to the recurrence
    assert np.array_equal(W.legendre_table(x01, 29), tab)


def test_gamma_tet_against_40_digit_brute_force():
    g = load("gamma_tiny.json")
    L, R = g["L"], g["R"]
    ws = W.lspace(L, R)
    G = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"])
    for a, row in enumerate(g["gamma_upper"]):
        for off, s in enumerate(row):
            b = a + off
            ref = float(mp.mpf(s))
            assert G[a, b] == pytest.approx(ref, rel=1e-13, abs=1e-13 * abs(G[a, a]))
            assert G[b, a] == G[a, b]


def test_gamma_gl_equals_tetrapyd_sum():
    # the Gauss-Legendre factorisation is an independent evaluation of the same sum
    for L, R, mod in [(60, 4, False), (97, 3, True), (150, 5, True)]:
        ws = W.lspace(L, R, mod)
        A = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"])
        B = O.gamma_gl(2, L, ws["q"], ws["v"], ws["order"])
        d = np.sqrt(np.outer(np.diag(A), np.diag(A)))
        assert np.abs(A - B).max() / d.max() < 1e-12
        assert (np.abs(A - B) / d).max() < 1e-10


def test_gamma_gl_needs_enough_nodes():
    # a plausible mistake (too few nodes) must show: npt < (3L+1)/2 is inexact
    ws = W.lspace(60, 3)
    A = O.gamma_gl(2, 60, ws["q"], ws["v"], ws["order"])
    B = O.gamma_gl(2, 60, ws["q"], ws["v"], ws["order"], npt=60)
    assert np.abs(A - B).max() / np.abs(A).max() > 1e-8


def test_constant_mode_equals_weight_sum():
    L = 80
    ws = W.lspace(L, 2)
    G = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"])
    pts = O.walk_XXX(2, L)
    i, j, k = pts.T
    v = ws["v"]
    w = np.array([O.lib().or_permsix(*p) * O.lib().or_h2_exact(*p) for p in pts.tolist()]) * v[i] * v[j] * v[k]
    assert G[0, 0] == pytest.approx(w.sum(), rel=1e-13)


def test_orthonormalise_invariants():
    rng = np.random.default_rng(3)
    X = rng.standard_normal((60, 40))
    G = X.T @ X + 40 * np.eye(40)
    st, ortho, lam, norm = O.orthonormalise(G)
    assert st == 0
    assert np.allclose(ortho, ortho.T) and np.allclose(lam, lam.T)
    F = np.tril(ortho)
    assert np.allclose(F @ F.T, G, rtol=1e-12, atol=1e-12 * np.abs(G).max())
    Lam = np.tril(lam)
    assert np.allclose(Lam @ G @ Lam.T, np.eye(40), atol=1e-11)
    c = rng.standard_normal(40)
    beta = O.cholesky_solve(ortho, G @ c)
    assert np.allclose(beta, c, rtol=1e-10, atol=1e-10)
    # beta^R = lambda alpha equals (D L)^T c
    assert np.allclose(Lam @ (G @ c), F.T @ c, rtol=1e-10, atol=1e-10)


def test_orthonormalise_reports_not_spd():
    G = np.ones((5, 5)) + np.diag([1.0, 1.0, 0.0, 1.0, 1.0])
    G[2, :] = G[1, :]
    G[:, 2] = G[:, 1]
    st, *_ = O.orthonormalise(G)
    assert st == -3


def test_calculate_weight_sum_rule():
    # cube fractions times ordered multiplicity sum to the volume of
    # {k in [0,M]^3 : triangle} = M^3 / 2 (in grid units)
    for M in (10, 25):
        pts = O.walk_prim(0, M)
        s = sum(O.lib().or_calculate_weight(0, M, *p) for p in pts.tolist())
        assert s == pytest.approx(M ** 3 / 2, rel=2e-2)


def test_kgl_gram_symmetry_and_limits():
    ks = W.kspace(14, 3)
    lim = O.kgl_limits(ks["kn"])
    kn = ks["kn"]
    for a in range(14):
        for b in range(14):
            c = lim[a, b]
            assert kn[c] <= kn[a] + kn[b] and (c == 13 or kn[c + 1] > kn[a] + kn[b])
    G = O.gram_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"])
    # brute force over the full masked cube
    Q = np.array([[O.lib().or_pijk(ks["q"].ctypes.data_as(__import__("ctypes").c_void_p), 3,
                                   ks["order"].ctypes.data_as(__import__("ctypes").c_void_p), m, a, b, c)
                   for m in range(ks["order"].shape[0])]
                  for a in range(14) for b in range(14) for c in range(14)])
    a, b, c = np.meshgrid(np.arange(14), np.arange(14), np.arange(14), indexing="ij")
    ka, kb, kc = kn[a].ravel(), kn[b].ravel(), kn[c].ravel()
    mask = (np.maximum(np.maximum(ka, kb), kc) * 2 <= ka + kb + kc)
    W3 = (ks["kw"][a] * ks["kw"][b] * ks["kw"][c]).ravel() / (ka + kb + kc) * mask
    Gb = (Q * W3[:, None]).T @ Q
    assert np.abs(G - Gb).max() / np.abs(Gb).max() < 1e-13


def test_projection_derivatives_vs_fd():
    ks = W.kspace(16, 3)
    om, ph, h = 150.0, 0.7, 1e-4
    a, dom, dph = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om, ph)
    ap, _, _ = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om + h, ph)
    am, _, _ = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om - h, ph)
    assert np.allclose((ap - am) / (2 * h), dom, rtol=1e-6, atol=1e-9 * np.abs(dom).max())
    bp, _, _ = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om, ph + h)
    bm, _, _ = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], om, ph - h)
    assert np.allclose((bp - bm) / (2 * h), dph, rtol=1e-6, atol=1e-9 * np.abs(dph).max())


def test_feature_phase_convention():
    # calculate_shape.c:933-946: phase = 2 * 3.1415927 * a2 / 100
    s = O.lib().or_shape3_feature(0.01, 0.02, 0.025, 150.0, 11.0)
    assert s == pytest.approx(np.sin(150.0 * 0.055 + 2 * 3.1415927 * 11.0 * 1e-2), rel=1e-12)
    assert O.lib().or_feature(0.0, 0.1, 0.1, 1.0, 0.0) == 0.0


def _sym(n, seed, cond=None):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    ev = rng.standard_normal(n) if cond is None else np.geomspace(1.0, 1.0 / cond, n) * rng.choice([-1, 1], n)
    return (q * ev) @ q.T


@pytest.mark.parametrize("n,seed", [(1, 0), (2, 1), (7, 2), (50, 3), (101, 4)])
def test_eigen_sym_against_lapack(n, seed):
    """Jacobi eigenvalues against LAPACK (numpy eigh, an independent
    algorithm); eigenvectors up to sign against LAPACK's, and A V = V diag(e)."""
    A = _sym(n, seed)
    A = 0.5 * (A + A.T)
    st, e, V = O.eigen_sym(A)
    assert st >= 0
    ref, W = np.linalg.eigh(A)
    ref, W = ref[::-1], W[:, ::-1]
    nrm = np.abs(ref).max()
    assert np.abs(e - ref).max() < 1e-13 * max(nrm, 1.0) * max(n, 1)
    assert np.abs(A @ V - V * e).max() < 1e-12 * max(nrm, 1.0)
    assert np.abs(V.T @ V - np.eye(n)).max() < 1e-12
    gaps = np.abs(np.diff(ref))
    for k in range(n):
        sep = min(gaps[k - 1] if k > 0 else np.inf, gaps[k] if k < n - 1 else np.inf)
        if sep > 1e-3 * nrm:
            assert abs(abs(V[:, k] @ W[:, k]) - 1.0) < 1e-10
        i = np.argmax(np.abs(V[:, k]))
        assert V[i, k] > 0          # sign convention: largest component positive


def test_eigen_sym_graded_spectrum():
    """Eigenvalues spread over 12 decades with both signs: each to a relative
    accuracy set by the largest, as a backward-stable method guarantees."""
    A = _sym(60, 9, cond=1e12)
    st, e, V = O.eigen_sym(A)
    ref = np.linalg.eigvalsh(A)[::-1]
    assert np.abs(e - ref).max() < 1e-13


def test_ortho_late_pca_identities():
    """PCA branch (forced, pca_flag = 0): lambda G lambda^T = I on the r
    positive directions and lambda lambdainv = I_r, zero outside."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((40, 25))
    G = X.T @ X + 1e-3 * np.eye(25)
    r, lam, linv, norm, ev = O.ortho_late(G, pca_flag=0)
    assert r == 25
    assert np.allclose(lam @ G @ lam.T, np.eye(25), atol=1e-10)
    assert np.allclose(lam @ linv, np.eye(25), atol=1e-10)
    assert np.allclose(linv @ linv.T, G, rtol=0, atol=1e-10 * np.abs(G).max())


def test_ortho_late_pca_on_indefinite_gram():
    """A Gram with a negative eigenvalue takes the PCA branch even with
    pca_flag = 1; r counts the positive eigenvalues; rows of lambda at and
    beyond r are zero."""
    rng = np.random.default_rng(11)
    q, _ = np.linalg.qr(rng.standard_normal((30, 30)))
    A = (q * np.r_[np.linspace(2.0, 0.5, 27), -0.05, -0.1, -0.2]) @ q.T
    assert np.diag(A).min() > 0
    d = rng.uniform(0.5, 3.0, 30)
    G = A * np.outer(d, d)
    w = np.linalg.eigvalsh(G / np.sqrt(np.outer(np.diag(G), np.diag(G))))
    assert w.min() < 0 < w.max()
    r, lam, linv, norm, ev = O.ortho_late(G, pca_flag=1)
    assert r == int((w > 0).sum())
    assert np.all(lam[r:] == 0) and np.all(linv[:, r:] == 0)
    P = lam[:r] @ G @ lam[:r].T
    assert np.allclose(P, np.eye(r), atol=1e-9)


def test_ortho_late_cholesky_branch():
    """SPD Gram with pca_flag = 1: lambdainv = D L (lower), lambda = its inverse."""
    rng = np.random.default_rng(6)
    X = rng.standard_normal((60, 20))
    G = X.T @ X
    r, lam, linv, norm, ev = O.ortho_late(G, pca_flag=1)
    assert r == 20
    assert np.allclose(np.triu(linv, 1), 0) and np.allclose(np.triu(lam, 1), 0)
    assert np.allclose(linv @ linv.T, G, rtol=1e-12, atol=1e-12 * np.abs(G).max())
    assert np.allclose(lam @ linv, np.eye(20), atol=1e-11)
