"""Independent pins of the oracle functions that golden vectors do not cover.

Each check comes from mathematics or from a second enumeration, not from the
oracle's own formula:

* calculate_weight (calculate_volume.c:12-132): every vertex of the tetrapyd
  {x <= y+z, y <= x+z, z <= x+y} clipped to one unit cell of the integer grid
  is a cell corner (the faces are planes through lattice points), so the
  region in a cell is the convex hull of its inside corners.  The weight of
  grid point p is then its ordered multiplicity times the exact integral of
  the trilinear hat function of p over the clipped tetrapyd, computed here in
  rational arithmetic over a triangulation of each cell's hull.  The cell-rule
  constants of the reference play no part in it.
* calculate_ortho / calculate_eigen on the uniform grid (orthogonalisation.c:
  14-86, calculate_eigen.c:86-145): a numpy sum over a separately generated
  point list with those exact weights.
* the Gauss-Legendre projection alpha: a numpy sum over the full masked cube
  (no ordering, no multiplicity), and alpha(omega = 0, phi = pi/2) = the
  Gram's column of the constant mode.
* get_next_ijk_bound / decompose_gamma_TTT_3D_mpi (calculate_gamma_3d.c:
  92-272): blocks end at the walk point a block size further on, and the
  thread ranges of every rank tile the walk.
* calculate_geometric (calculate_geometric.c:59-92): the formula evaluated in
  40-digit arithmetic.
"""
from fractions import Fraction
from itertools import permutations, product

import mpmath as mp
import numpy as np
import pytest

import pyoracle as O
from modalcmb import workloads as W

try:
    from scipy.spatial import ConvexHull
except ImportError:  # pragma: no cover
    ConvexHull = None


# ---- exact cell integrals ---------------------------------------------------

def _tet_hat_integral(V, c, lo):
    """Exact integral over the tetrahedron V (4 integer vertices) of the
    trilinear hat of grid point c restricted to the cell with low corner lo."""
    def lin(a, x):   # the hat's factor along axis a on this cell (affine)
        return (x[a] - lo[a]) if c[a] == lo[a] + 1 else (lo[a] + 1 - x[a])
    M = [[V[t][a] - V[0][a] for a in range(3)] for t in (1, 2, 3)]
    det = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1])
           - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0])
           + M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]))
    vol = Fraction(abs(det), 6)
    if vol == 0:
        return Fraction(0)
    # product of three affine functions = cubic form in barycentrics; over a
    # tetrahedron, int l_i l_j l_k = vol * {1/20, 1/60, 1/120} for 1, 2, 3
    # distinct indices
    s = Fraction(0)
    for i, j, k in product(range(4), repeat=3):
        d = len({i, j, k})
        m = Fraction(1, 20) if d == 1 else (Fraction(1, 60) if d == 2 else Fraction(1, 120))
        s += lin(0, V[i]) * lin(1, V[j]) * lin(2, V[k]) * m
    return vol * s


def exact_weight(mn, mx, i, j, k):
    def inside(p):
        a, b, c = p
        return min(p) >= mn and max(p) <= mx and a <= b + c and b <= a + c and c <= a + b
    tot = Fraction(0)
    for s in product((0, 1), repeat=3):
        lo = (i - 1 + s[0], j - 1 + s[1], k - 1 + s[2])
        S = [p for p in (tuple(lo[a] + d[a] for a in range(3)) for d in product((0, 1), repeat=3)) if inside(p)]
        if len(S) < 4:
            continue
        P = np.array(S, dtype=float)
        if np.linalg.matrix_rank(P[1:] - P[0]) < 3:
            continue
        hull = ConvexHull(P)
        for f in hull.simplices:   # fan from one hull vertex over the triangular facets
            tot += _tet_hat_integral([S[0]] + [S[t] for t in f], (i, j, k), lo)
    mult = 1 if i == k else (3 if (i == j or j == k) else 6)
    return mult * tot


@pytest.mark.skipif(ConvexHull is None, reason="scipy not importable")
@pytest.mark.parametrize("mn,mx", [(0, 6), (0, 9), (2, 9), (1, 12)])
def test_calculate_weight_is_exact_hat_integral(mn, mx):
    pts = O.walk_prim(mn, mx).tolist()
    for (i, j, k) in pts:
        a = O.lib().or_calculate_weight(mn, mx, i, j, k)
        b = float(exact_weight(mn, mx, i, j, k))
        assert a == pytest.approx(b, rel=0, abs=1e-15), (i, j, k)


# ---- uniform-grid Gram and alpha --------------------------------------------

def _modes_at(q, order, pts):
    """Q_n at points (P, 3): mean over the 6 permutations of the index slots."""
    Q = np.zeros((pts.shape[0], order.shape[0]))
    for perm in permutations(range(3)):
        Q += (q[pts[:, perm[0]]][:, order[:, 0]] * q[pts[:, perm[1]]][:, order[:, 1]]
              * q[pts[:, perm[2]]][:, order[:, 2]])
    return Q / 6.0


@pytest.mark.skipif(ConvexHull is None, reason="scipy not importable")
def test_uniform_gram_and_alpha_brute_force():
    mx, kmax, R = 10, 0.2, 3
    grid = np.arange(mx + 1) / mx
    q = W.legendre_table(grid, R - 1)
    order = W.order_lex(R)
    pts = np.array([(i, j, k) for i in range(mx + 1) for j in range(i, mx + 1) for k in range(j, mx + 1)
                    if k <= i + j + 2 and (i < mx or (j == mx and k == mx))])
    dk = kmax / mx
    w = np.array([float(exact_weight(0, mx, *p)) for p in pts.tolist()]) * dk ** 3
    ks = dk * pts.sum(1)
    inv = np.where(ks > 1e-10, 1.0 / np.where(ks > 1e-10, ks, 1.0), ks)
    x = w * inv
    Q = _modes_at(q, order, pts)
    G = (Q * x[:, None]).T @ Q
    iu = np.triu_indices(order.shape[0])
    packed = O.calculate_ortho_uniform(mx, kmax, q, order)
    assert np.abs(packed - G[iu]).max() < 1e-14 * np.abs(G).max()
    # feature shape (calculate_shape.c:933-946): shape3 = sin(a1 K + 2 pi' a2 / 100)
    # with pi' = 3.1415927, 0 where k1 k2 k3 <= 1e-15
    a1, a2 = 150.0, 11.0
    k = dk * pts
    s3 = np.where(k.prod(1) > 1e-15, np.sin(a1 * k.sum(1) + 2.0 * 3.1415927 * a2 * 1e-2), 0.0)
    alpha = (Q * (x * s3)[:, None]).sum(0)
    ref = O.calculate_eigen_uniform(mx, kmax, q, order, a1, a2)
    assert np.abs(ref - alpha).max() < 1e-13 * np.abs(alpha).max()


# ---- Gauss-Legendre projection ------------------------------------------------

def test_project_kgl_full_cube_brute_force():
    ks = W.kspace(14, 4)
    kn, kw, q, order = ks["kn"], ks["kw"], ks["q"], ks["order"]
    N = len(kn)
    a, b, c = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    pts = np.stack([a.ravel(), b.ravel(), c.ravel()], 1)
    kk = kn[pts]
    srt = np.sort(kk, 1)
    keep = srt[:, 2] <= srt[:, 0] + srt[:, 1]
    pts, kk = pts[keep], kk[keep]
    K = kk.sum(1)
    W3 = kw[pts].prod(1) / K
    Q = _modes_at(q, order, pts)
    for om, ph in [(150.0, 0.7), (37.5, 3.3)]:
        ref = (Q * (W3 * np.sin(om * K + ph))[:, None]).sum(0)
        dom = (Q * (W3 * K * np.cos(om * K + ph))[:, None]).sum(0)
        dph = (Q * (W3 * np.cos(om * K + ph))[:, None]).sum(0)
        al, d1, d2 = O.project_kgl(kn, kw, q, order, om, ph)
        assert np.abs(al - ref).max() < 1e-13 * np.abs(ref).max()
        assert np.abs(d1 - dom).max() < 1e-13 * np.abs(dom).max()
        assert np.abs(d2 - dph).max() < 1e-13 * np.abs(dph).max()


def test_project_kgl_constant_shape_is_gram_column():
    ks = W.kspace(20, 5)
    G = O.gram_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"])
    assert tuple(ks["order"][0]) == (0, 0, 0)   # q_0 = P_0 = 1, so Q_0 = 1
    al, _, _ = O.project_kgl(ks["kn"], ks["kw"], ks["q"], ks["order"], 0.0, np.pi / 2)
    assert np.abs(al - G[:, 0]).max() < 1e-14 * np.abs(G[:, 0]).max()


# ---- block bounds and the 3D decomposition -----------------------------------

def _walk(lsize):
    return [tuple(p) for p in O.walk_XXX(2, lsize - 1).tolist()]


def test_next_ijk_bound_against_walk():
    rng = np.random.default_rng(5)
    for lsize in (20, 33, 48):
        Wk = _walk(lsize)
        N = len(Wk)
        sent = (lsize, lsize, lsize)
        at = lambda x: Wk[x] if x < N else sent
        for _ in range(200):
            s = int(rng.integers(0, N))
            mb = int(rng.integers(0, 300))
            f = int(rng.integers(s, N + 1))        # final bound: a later walk point or the end
            fin = at(f)
            got = O.next_ijk_bound(lsize, *Wk[s], *fin, mb)
            e = min(s + mb, f)
            assert got == (*at(e), e - s), (lsize, s, mb, f)


@pytest.mark.parametrize("lsize", [21, 40])
def test_decompose_gamma_3d_tiles_walk(lsize):
    Wk = _walk(lsize)
    N = len(Wk)
    sent = (lsize, lsize, lsize)
    at = lambda x: Wk[x] if x < N else sent
    for numranks, nthreads, ws in [(1, 1, 0.0), (1, 6, 0.0), (3, 4, 0.0), (4, 3, 0.25), (7, 2, 0.0)]:
        blk, left = divmod(N, numranks)
        covered = []
        start = 0
        for r in range(numranks):
            work = blk + (r < left)
            host = int((1.0 - ws) * work)
            b6 = O.decompose_gamma_TTT_3D(lsize, r, numranks, ws, nthreads)
            tb, tl = divmod(host, nthreads)
            s = start
            for t in range(nthreads):
                w = tb + (t < tl)
                assert tuple(b6[t, :3]) == at(s) and tuple(b6[t, 3:]) == at(s + w), (r, t)
                covered.extend(range(s, s + w))
                s += w
            start += work
        if ws == 0.0:
            assert covered == list(range(N))


# ---- asymptotic h^2 ------------------------------------------------------------

def test_geometric_asym_against_40_digit_evaluation():
    mp.mp.dps = 40
    for l1, l2, l3 in [(2, 2, 2), (2, 3, 5), (10, 20, 30), (100, 200, 300), (7, 1500, 1501),
                       (2500, 2500, 2500), (1, 1, 2), (0, 5, 5)]:
        a = O.lib().or_geometric_asym(l1, l2, l3)
        if (l1 + l2 + l3) % 2 or l1 + l2 < l3 or l2 + l3 < l1 or l3 + l1 < l2:
            assert a == 0.0
            continue
        L = mp.mpf(l1 + l2 + l3)
        t = [mp.mpf(l2 + l3 - l1), mp.mpf(l3 + l1 - l2), mp.mpf(l1 + l2 - l3)]
        p1 = (2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (2 * mp.pi ** 2 * (L + 1))
        p2 = (L + mp.mpf(1) / 3) / ((t[0] + mp.mpf(1) / 3) * (t[1] + mp.mpf(1) / 3) * (t[2] + mp.mpf(1) / 3))
        p3 = mp.sqrt((t[0] + mp.mpf(1) / 6) * (t[1] + mp.mpf(1) / 6) * (t[2] + mp.mpf(1) / 6) / (L + mp.mpf(1) / 6))
        ref = float(p1 * p2 * p3)
        assert a == pytest.approx(ref, rel=1e-14, abs=0), (l1, l2, l3)
