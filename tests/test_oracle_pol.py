"""Pins of the oracle's polarisation parts (no GPU).

* The XXY walk: where it is well posed (lminX = lminY and lmaxX = lmaxY,
  both even) its sequence is the brute-force enumeration of the mixed tetrapyd
  {i <= j in X, j - i <= k <= i + j, k in Y, even sum} in (i, j, k) order;
  its quirks elsewhere are held by single steps traced by hand from
  tetrapyd_tools.c:299-414; decompose_tetrapyd_XXY's slices tile the walk.
* The Gauss-Legendre Gram of the TTE and TEE components (calculate_orthol's
  1/(16 pi) formulas) equals the direct sum over the mixed tetrapyd with the
  exact h^2 and multiplicity 2 for unequal pair legs -- an independent
  algorithm; TTT and EEE equal the direct tetrapyd Gram of their field.
"""
import numpy as np
import pytest

import pyoracle as O
from modalcmb import workloads as W


def brute_mixed(lminX, lminY, lmaxX, lmaxY):
    pts = []
    for i in range(lminX, lmaxX + 1):
        for j in range(i, lmaxX + 1):
            for k in range(max(j - i, lminY), min(i + j, lmaxY) + 1):
                if (i + j + k) % 2 == 0:
                    pts.append((i, j, k))
    return np.array(pts, dtype=np.int32)


@pytest.mark.parametrize("lmin,lmax", [(2, 4), (2, 10), (2, 20), (2, 32), (4, 26), (6, 42)])
def test_xxy_walk_is_the_mixed_tetrapyd_where_well_posed(lmin, lmax):
    assert np.array_equal(O.walk_XXY(lmin, lmin, lmax, lmax), brute_mixed(lmin, lmin, lmax, lmax))


def test_xxy_count_of_one_point_is_two():
    # the count loop (tetrapyd_tools.c:243-251) ends with count++ after the
    # step that reaches the last point, so a one-point range counts 2
    assert O.lib().or_count_tetrapyd_XXY(2, 2, 2, 2) == 2


# (lminX, lminY, lmaxX, lmaxY, point) -> (more, next point), traced by hand
# through get_ijk_next_XXY (tetrapyd_tools.c:299-337)
STEPS = [
    # ordinary k step, then a j step whose k restarts at max(j - i, lminY), parity-bumped
    ((2, 2, 20, 20, (2, 5, 3)), (1, (2, 5, 5))),
    ((2, 2, 20, 20, (2, 5, 7)), (1, (2, 6, 4))),
    # j at lmaxX and k past lmaxY: i advances, k = lminY + (2i % 2) = lminY
    ((2, 2, 20, 20, (2, 20, 20)), (1, (3, 3, 2))),
    # odd lminY: the restart keeps k = lminY, an odd-sum point
    ((2, 3, 20, 20, (2, 20, 20)), (1, (3, 3, 3))),
    # lmaxY < j - i: the step still yields (i, j, j - i) above lmaxY
    ((2, 2, 30, 17, (2, 19, 17)), (1, (2, 20, 18))),
    # the end: i + j + k + 1 >= 2 lmaxX + lmaxY sets the last point
    ((2, 2, 20, 20, (20, 20, 18)), (0, (20, 20, 20))),
    ((2, 2, 21, 21, (21, 21, 18)), (0, (21, 21, 20))),
    ((2, 2, 20, 25, (19, 20, 23)), (0, (20, 20, 24))),
    # odd lmaxY: (lmaxX - 1, lmaxX, lmaxY) has the last point's sum and ends
    # the walk, skipping it and the row i = lmaxX
    ((2, 2, 9, 9, (8, 9, 7)), (0, (9, 9, 8))),
    # j = lmaxX, k past lmaxY > lmaxX: the next i restarts at k = lminY
    ((2, 2, 20, 25, (19, 20, 25)), (1, (20, 20, 2))),
]


def test_xxy_walk_odd_lmax_ends_early():
    w = O.walk_XXY(2, 2, 9, 9)
    assert len(w) == 119 and tuple(w[-2]) == (8, 9, 7) and tuple(w[-1]) == (9, 9, 8)
    full = brute_mixed(2, 2, 9, 9)
    assert len(full) == 123 and np.array_equal(w[:-1], full[:118])


@pytest.mark.parametrize("args,expect", STEPS)
def test_xxy_step_known_answers(args, expect):
    *lims, p = args
    assert O.next_XXY(*lims, p) == expect


# check_ijk_min_XXY / check_ijk_max_XXY (tetrapyd_tools.c:339-414), by hand
MINS = [
    ((2, 2, (1, 0, 0)), (2, 2, 2)),        # clamped, already valid
    ((2, 2, (9, 4, 2)), (4, 6, 2)),        # swap i, j; j > i + k: j = i + k
    ((2, 2, (3, 4, 4)), (3, 4, 5)),        # odd sum: k + 1
    ((2, 2, (3, 4, 9)), (3, 4, 7)),        # k > i + j >= lminY: k = i + j
    ((2, 10, (2, 2, 0)), (5, 5, 10)),      # i + j < lminY > 2 lminX: (lminY/2, lminY/2 + lminY%2, lminY)
    ((4, 7, (4, 4, 0)), (4, 4, 8)),        # k clamped to lminY, odd sum: k + 1
]
MAXS = [
    ((20, 20, (30, 25, 25)), (20, 20, 20)),
    ((20, 20, (9, 4, 2)), (4, 6, 2)),      # swap, j > i + k: j = i + k
    ((20, 20, (3, 4, 4)), (3, 4, 3)),      # odd sum: k - 1
    ((20, 15, (3, 4, 9)), (3, 4, 7)),      # k > i + j: k = i + j
]


@pytest.mark.parametrize("args,expect", MINS)
def test_xxy_check_min_known_answers(args, expect):
    assert O.check_min_XXY(*args) == expect


@pytest.mark.parametrize("args,expect", MAXS)
def test_xxy_check_max_known_answers(args, expect):
    assert O.check_max_XXY(*args) == expect


@pytest.mark.parametrize("lims,rank,nranks,nthreads",
                         [((2, 2, 20, 20), 0, 1, 4), ((2, 2, 25, 25), 1, 3, 5), ((2, 3, 17, 22), 2, 4, 3),
                          ((3, 2, 21, 26), 0, 2, 7)])
def test_xxy_decompose_slices_tile_the_walk(lims, rank, nranks, nthreads):
    walk = O.walk_XXY(*lims)
    n = len(walk)
    per_rank = [n // nranks + (r < n % nranks) for r in range(nranks)]
    off = sum(per_rank[:rank])
    blk = O.decompose_XXY(rank, nranks, nthreads, *lims)
    assert blk[:, 3].sum() == per_rank[rank]
    for t in range(nthreads):
        assert tuple(blk[t, :3]) == tuple(walk[min(off, n - 1)])
        off += blk[t, 3]


def field(lmin, lmax, R, seed):
    """a Legendre basis on [lmin, lmax] with seeded column scales, and a
    smooth per-leg weight"""
    rng = np.random.default_rng(seed)
    ls = np.arange(lmax + 1.0)
    q = np.zeros((lmax + 1, R))
    q[lmin:] = W.legendre_table((ls[lmin:] - lmin) / (lmax - lmin), R - 1) * (1 + 0.3 * rng.standard_normal(R))
    v = (1 + 0.1 * np.sin(ls / 7 + seed)) / (1 + 0.01 * ls) ** 2
    return (lmin, lmax, q, v)


def pol_tables(RT, RE):
    tte = np.array([(a, b, c) for a in range(RT) for b in range(a, RT) for c in range(RE)], dtype=np.int32)
    tee = np.array([(a, b, c) for a in range(RT) for b in range(RE) for c in range(b, RE)], dtype=np.int32)
    return tte, tee


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


@pytest.mark.parametrize("Tr,Er,RT,RE", [((2, 24), (2, 24), 4, 3), ((2, 30), (3, 22), 3, 4), ((5, 20), (2, 28), 4, 4)])
def test_gl_pol_equals_direct_mixed_sum(Tr, Er, RT, RE):
    T = field(*Tr, RT, 1)
    E = field(*Er, RE, 2)
    tte, tee = pol_tables(RT, RE)
    for pol, order in ((1, tte), (2, tee)):
        assert relmax(O.gamma_gl_pol(pol, T, E, order), O.gamma_xxy_direct(pol, T, E, order)) < 1e-12
    eee = W.order_lex(RE).astype(np.int32)
    assert relmax(O.gamma_gl_pol(3, T, E, eee), O.gamma_tet(E[0], E[1], E[2], E[3], eee)) < 1e-12
    ttt = W.order_lex(RT).astype(np.int32)
    assert relmax(O.gamma_gl_pol(0, T, E, ttt), O.gamma_tet(T[0], T[1], T[2], T[3], ttt)) < 1e-12


def test_direct_mixed_sum_swaps_fields():
    """TEE over (T, E) is TTE over (E, T) with the slots rotated: the pair
    legs of TEE are E legs"""
    T = field(2, 22, 3, 5)
    E = field(2, 26, 4, 6)
    tte, tee = pol_tables(3, 4)
    # TEE modes (t, e1, e2) = TTE modes of the swapped fields (e1, e2, t)
    g_tee = O.gamma_xxy_direct(2, T, E, tee)
    g_swapped = O.gamma_xxy_direct(1, E, T, tee[:, [1, 2, 0]])
    assert relmax(g_tee, g_swapped) < 1e-13


# ---- the polarised cross-Gamma ------------------------------------------------

FIELDS_OF = {0: (0, 0, 0), 1: (0, 0, 1), 2: (0, 1, 1), 3: (1, 1, 1)}


def spectra(lsize):
    """C_TT, C_TE, C_EE, N_TT, N_TE, N_EE, b_TT, b_TE, b_EE with
    s1 s4 - s3^2 > 0 at every l"""
    ls = np.arange(lsize, dtype=np.float64)
    ctt = np.where(ls > 0, 1000.0 / np.maximum(ls * (ls + 1), 1.0), 0.0)
    cee = 0.05 * ctt * (1 + 0.2 * np.cos(ls / 9))
    cte = 0.4 * np.sqrt(ctt * cee) * np.sin(ls / 5)
    ntt, nte, nee = 0.02 + 0 * ls, 0.003 * np.cos(ls / 4), 0.03 + 0 * ls
    btt, bte, bee = (np.exp(-ls * (ls + 1) * s) for s in (2e-4, 3e-4, 4e-4))
    return np.array([ctt, cte, cee, ntt, nte, nee, btt, bte, bee])


def separable_qtilde(T, E, spec, np_, nx, fsky_ratio, seed):
    """q-tilde with x dependence g(x) only, built so that TTw qt_T = g (2l+1)
    v_T b^T_p over the T range and EEw (s1 qt_E - s2 qt_T) = g (2l+1) v_E b^E_p
    over the E range (the weights of calculate_gamma.c:73-160).  g is scaled by
    1e-9 on the first 10 samples so that the x < 10 rule keeps them."""
    rng = np.random.default_rng(seed)
    (tl, th, qT, vT), (el, eh, qE, vE) = T, E
    lsq = spec.shape[1]
    CTT, CTE, CEE, NTT, NTE, NEE, BTT, BTE, BEE = spec
    x = np.sort(rng.uniform(0.5, 3.0, nx))
    g = np.cos(x) + 1.5
    g[:10] *= 1e-9
    ls = np.arange(lsq, dtype=np.float64)
    s1 = BTT ** 2 * CTT + NTT
    s2 = np.sqrt(fsky_ratio) * BTE ** 2 * CTE + NTE
    s3 = BTT * BEE * CTE + NTE
    s4 = BEE ** 2 * CEE + NEE
    with np.errstate(divide="ignore", invalid="ignore"):
        ttw = (2 * ls + 1) ** (5 / 6) * BTT / np.sqrt(s1)
        eew = (2 * ls + 1) ** (5 / 6) * BEE / np.sqrt(s1 * (s1 * s4 - s3 ** 2))
    qtT = np.zeros((np_, lsq, nx))
    qtE = np.zeros((np_, lsq, nx))
    for l in range(max(tl, 2), th + 1):
        qtT[:, l, :] = np.outer((2 * l + 1) * vT[l] * qT[l, :np_] / ttw[l], g)
    for l in range(max(el, 2), eh + 1):
        comb = np.outer((2 * l + 1) * vE[l] * qE[l, :np_] / eew[l], g)
        qtE[:, l, :] = (comb + s2[l] * qtT[:, l, :]) / s1[l]
    return x, g, qtT, qtE


def cube_sum(T, E, pol, L):
    """D[a1, q1, a2, q2, a3, q3] = sum over the full cube (legs in their
    fields' ranges) of h^2 prod_b v(l_b) b_{a_b}(l_b) b_{q_b}(l_b), exact h^2"""
    H = np.zeros((L, L, L))
    for i in range(L):
        for j in range(L):
            for k in range(L):
                if (i + j + k) % 2 == 0 and abs(i - j) <= k <= i + j:
                    H[i, j, k] = O.lib().or_h2_exact(i, j, k)
    A = []
    for f in FIELDS_OF[pol]:
        lo, hi, q, v = T if f == 0 else E
        a = np.zeros((L, q.shape[1], q.shape[1]))
        for l in range(lo, hi + 1):
            a[l] = v[l] * np.outer(q[l], q[l])
        A.append(a)
    return np.einsum("ijk,iab,jcd,kef->abcdef", H, A[0], A[1], A[2], optimize=True)


@pytest.mark.parametrize("pol", [0, 1, 2, 3])
@pytest.mark.parametrize("spline", [0, 1])
def test_cross_gl_pol_reduces_to_mixed_cube_sums(pol, spline):
    R = 3
    T = field(2, 22, R, 11)
    E = field(3, 26, R, 12)
    L = 27
    spec = spectra(L)
    fr, dphi = 0.8, 1.5714e-8
    x, g, qtT, qtE = separable_qtilde(T, E, spec, R, 24, fr, 13)
    late = W.order_lex(R).astype(np.int32)            # any slot assignment of indices < R
    rng = np.random.default_rng(14)
    prim = W.order_lex(R).astype(np.int32)[rng.permutation(10)]
    G = O.gamma_cross_gl_pol(pol, T[:3], E[:3], late, qtT, qtE, x, prim, spec, fr, dphi, spline=spline)
    y = x ** 2 * g ** 3
    Ix = O.spline_integral(x, y, "natural" if spline else "fmm")
    D = cube_sum(T, E, pol, L)
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    ref = np.empty_like(G)
    for m, (q1, q2, q3) in enumerate(late):
        for c, pn in enumerate(prim):
            ref[m, c] = sum(D[pn[s[0]], q1, pn[s[1]], q2, pn[s[2]], q3] for s in perms)
    ref *= dphi ** 2 * Ix
    assert relmax(G, ref) < 1e-11


def test_cross_gl_pol_ttt_is_the_tt_cross_gamma():
    """component 0 with no E field in use equals the TT-only restatement"""
    R = 3
    T = field(2, 22, R, 21)
    E = field(2, 22, R, 22)
    spec = spectra(23)
    rng = np.random.default_rng(23)
    qtT = rng.standard_normal((R, 23, 12))
    qtE = rng.standard_normal((R, 23, 12))
    x = np.sort(rng.uniform(0.2, 3.0, 12))
    late = W.order_lex(R).astype(np.int32)
    prim = late[rng.permutation(len(late))]
    for spline in (0, 1):
        G = O.gamma_cross_gl_pol(0, T[:3], E[:3], late, qtT, qtE, x, prim, spec, 0.7, 2e-8, spline=spline)
        ref = O.gamma_cross_gl(2, 22, T[2], late, qtT, x, prim, spec[0], spec[3], spec[6], 2e-8, spline=spline)
        assert relmax(G, ref) < 1e-14
