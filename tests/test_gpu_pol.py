"""GPU polarised Gauss-Legendre Gram (calculate_orthol result[0..3]) through
the C ABI against the oracle: all four components against the oracle's
restatement of calculate_orthol, TTE / TEE also against the direct sum over
the mixed tetrapyd; the reference-named shims (calculate_orthol,
find_perm_late_TTE/TEE/EEE, plijk_TTE/TEE/EEE) over a bound context."""
import ctypes as C
import os

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def no_E_after(ctx):
    """the session context leaves each test without an E setup"""
    yield
    ctx.lspace_clear_E()


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def field(lmin, lmax, R, seed):
    rng = np.random.default_rng(seed)
    ls = np.arange(lmax + 1.0)
    q = np.zeros((lmax + 1, R))
    q[lmin:] = W.legendre_table((ls[lmin:] - lmin) / (lmax - lmin), R - 1) * (1 + 0.3 * rng.standard_normal(R))
    v = (1 + 0.1 * np.sin(ls / 7 + seed)) / (1 + 0.01 * ls) ** 2
    return (lmin, lmax, q, v)


def tables(RT, RE, n=None, seed=0):
    """TTE (T, T, E), TEE (T, E, E) and EEE tables of one row count n (the
    reference reads terms_late rows of each order file)"""
    tte = np.array([(a, b, c) for a in range(RT) for b in range(a, RT) for c in range(RE)], dtype=np.int32)
    tee = np.array([(a, b, c) for a in range(RT) for b in range(RE) for c in range(b, RE)], dtype=np.int32)
    eee = W.order_lex(RE).astype(np.int32)
    n = n or min(len(tte), len(tee), len(eee))
    rng = np.random.default_rng(seed)
    return [t[np.sort(rng.permutation(len(t))[:n])] for t in (tte, tee, eee)]


def setup(ctx, T, E, RT, RE, n=None):
    ttt = W.order_lex(RT).astype(np.int32)
    tte, tee, eee = tables(RT, RE, n)
    if n is not None:
        ttt = ttt[:n]
    ctx.lspace_set(T[0], T[1], T[2], T[3], ttt, h2_mode=modalcmb.H2_ASYMPTOTIC)
    ctx.lspace_set_E(E[0], E[1], E[2], E[3], tte, tee, eee)
    return ttt, tte, tee, eee


@pytest.mark.parametrize("Tr,Er,RT,RE", [((2, 40), (2, 36), 5, 4), ((2, 60), (10, 50), 4, 5), ((7, 33), (2, 64), 6, 3)])
def test_gl_pol_components_vs_oracle(ctx, Tr, Er, RT, RE):
    T, E = field(*Tr, RT, 1), field(*Er, RE, 2)
    ttt, tte, tee, eee = setup(ctx, T, E, RT, RE)
    npt = 3 * max(Tr[1], Er[1]) // 2 + 1
    for pol, order in ((0, ttt), (1, tte), (2, tee), (3, eee)):
        G = ctx.gamma_gl_pol(pol)
        assert G.shape == (len(order), len(order)) and np.array_equal(G, G.T)
        ref = O.gamma_gl_pol(pol, T, E, order, npt=npt)
        assert relmax(G, ref) < 1e-12, pol
        if pol in (1, 2):
            assert relmax(G, O.gamma_xxy_direct(pol, T, E, order)) < 1e-12
    # TTT on its own node count is the auto GL Gram
    assert relmax(ctx.gamma_gl_pol(0, 3 * Tr[1] // 2 + 1), ctx.gamma_gl()) < 1e-14


def test_gl_eee_compressed_equals_direct_node_sum(ctx):
    """A Legendre-type E basis at lmax 2500 with 30 columns and all 4960 EEE
    modes: the EEE component runs over the compressed pair basis (rank 2R - 1
    = 59) and equals the direct node sum (MCMB_GL_PATH=direct) to rounding;
    the E basis has its own cached factorisation beside the T basis's."""
    T, E = field(2, 2500, 30, 5), field(2, 2500, 30, 6)
    setup(ctx, T, E, 30, 30, n=4960)
    G = ctx.gamma_gl_pol(modalcmb.POL_EEE)
    assert ctx.last_basis_rank() == 59
    os.environ["MCMB_GL_PATH"] = "direct"
    try:
        Gd = ctx.gamma_gl_pol(modalcmb.POL_EEE)
        assert ctx.last_basis_rank() == 0
    finally:
        del os.environ["MCMB_GL_PATH"]
    d = np.sqrt(np.outer(np.abs(np.diag(Gd)), np.abs(np.diag(Gd))))
    assert float((np.abs(G - Gd) / d).max()) < 1e-13
    assert np.array_equal(G, G.T)
    # the T basis's compression is kept apart: TTT and EEE alternate, each bit-stable
    Gt = ctx.gamma_gl()
    assert ctx.last_basis_rank() == 59
    assert np.array_equal(ctx.gamma_gl_pol(modalcmb.POL_EEE), G)
    assert np.array_equal(ctx.gamma_gl(), Gt)


@pytest.mark.parametrize("RT,RE,n", [(30, 30, 4960), (12, 20, 600), (20, 9, 160)])
def test_gl_mixed_compressed_equals_direct_node_sum(ctx, RT, RE, n):
    """TTE and TEE over the compressed T and E bases (ranks 2 RT - 1 and 2 RE
    - 1) against the direct node sum (MCMB_GL_PATH=direct): the full lmax
    2500 / 30-column size with 4960 modes, and unequal widths either way round
    (the field on two legs wider or narrower than the other)."""
    T, E = field(2, 2500, RT, 7), field(10, 2400, RE, 8)
    setup(ctx, T, E, RT, RE, n=n)
    for pol, rank in ((modalcmb.POL_TTE, 2 * RT - 1), (modalcmb.POL_TEE, 2 * RE - 1)):
        G = ctx.gamma_gl_pol(pol)
        assert ctx.last_basis_rank() == rank, pol
        os.environ["MCMB_GL_PATH"] = "direct"
        try:
            Gd = ctx.gamma_gl_pol(pol)
            assert ctx.last_basis_rank() == 0
        finally:
            del os.environ["MCMB_GL_PATH"]
        d = np.sqrt(np.outer(np.abs(np.diag(Gd)), np.abs(np.diag(Gd))))
        assert float((np.abs(G - Gd) / d).max()) < 1e-13, pol
        assert np.array_equal(G, G.T)
        assert np.array_equal(ctx.gamma_gl_pol(pol), G)


def test_gl_pol_lmax1000_vs_oracle(ctx):
    T, E = field(2, 1000, 12, 3), field(2, 900, 10, 4)
    ttt, tte, tee, eee = setup(ctx, T, E, 12, 10, n=200)
    for pol, order in ((1, tte), (2, tee), (3, eee)):
        assert relmax(ctx.gamma_gl_pol(pol), O.gamma_gl_pol(pol, T, E, order)) < 1e-12, pol


def test_gl_pol_errors(ctx):
    T, E = field(2, 30, 4, 5), field(2, 30, 3, 6)
    with modalcmb.Context(0) as c2:
        c2.lspace_set(T[0], T[1], T[2], T[3], W.order_lex(4))
        with pytest.raises(modalcmb.ModalError) as e:
            c2.nE = 1
            c2.gamma_gl_pol(modalcmb.POL_TTE)
        assert e.value.status == -7                      # no E setup
    ttt, tte, tee, eee = setup(ctx, T, E, 4, 3)
    with pytest.raises(modalcmb.ModalError) as e:
        ctx.gamma_gl_pol(modalcmb.POL_TEE, npt=40)       # 2 npt - 1 < 30 + 2 * 30
    assert e.value.status == -1
    # a T index beyond the l-space basis width is refused at launch
    bad = tte.copy()
    bad[0, 0] = 5
    ctx.lspace_set_E(E[0], E[1], E[2], E[3], bad, tee, eee)
    with pytest.raises(modalcmb.ModalError) as e:
        ctx.gamma_gl_pol(modalcmb.POL_TTE)
    assert e.value.status == -4


def test_legacy_calculate_orthol_pol(ctx):
    lib = modalcmb.load_library()
    RT, RE = 5, 5
    T, E = field(2, 44, RT, 7), field(2, 40, RE, 8)
    ttt = W.order_lex(RT).astype(np.int32)
    n = len(ttt)
    tte, tee, eee = tables(RT, RE, n=min(n, 40))
    ttt = ttt[:len(tte)]
    n = len(ttt)
    ctx.lspace_set(T[0], T[1], T[2], T[3], ttt, h2_mode=modalcmb.H2_ASYMPTOTIC)
    lib.mcmb_legacy_bind(ctx._h)
    try:
        res = (C.c_double * 4)(-1, -1, -1, -1)
        # no E setup: result[0] only
        assert lib.calculate_orthol(3, 5, res) == 0
        assert res[0] == pytest.approx(ctx.gamma_gl()[3, 5], rel=1e-14) and res[1] == -1
        lib.calculate_orthoMij_E(0, 0)
        assert lib.mcmb_legacy_status() == -7
        ctx.lspace_set_E(E[0], E[1], E[2], E[3], tte, tee, eee)
        G = [ctx.gamma_gl_pol(p) for p in range(4)]
        for m, c in ((0, 0), (3, 5), (n - 1, 2), (n - 1, n - 1)):
            assert lib.calculate_orthol(m, c, res) == 0
            assert [res[k] for k in range(4)] == [G[k][m, c] for k in range(4)]
        lib.calculate_orthoMij_E(0, 0)
        assert lib.mcmb_legacy_status() == 0
        p = [C.c_int() for _ in range(3)]
        for name, tab in (("find_perm_late_TTE", tte), ("find_perm_late_TEE", tee), ("find_perm_late_EEE", eee)):
            getattr(lib, name)(7, *(C.byref(x) for x in p))
            assert [x.value for x in p] == list(tab[7])
        # plijk_*: the reference's mixed products (orthogonalisation_late.c:646-715)
        qT, qE = T[2], E[2]
        for r in (0, 11, n - 1):
            a, b, c = tte[r]
            l1, l2, l3 = 9, 14, 21
            want = (qT[l1, a] * qT[l2, b] + qT[l1, b] * qT[l2, a]) * qE[l3, c] / 2
            assert lib.plijk_TTE(r, l1, l2, l3) == pytest.approx(want, rel=1e-15)
            a, b, c = tee[r]
            want = qT[l1, a] * (qE[l2, b] * qE[l3, c] + qE[l2, c] * qE[l3, b]) / 2
            assert lib.plijk_TEE(r, l1, l2, l3) == pytest.approx(want, rel=1e-15)
            a, b, c = eee[r]
            perms = [(a, b, c), (b, c, a), (c, a, b), (c, b, a), (b, a, c), (a, c, b)]
            want = sum(qE[l1, x] * qE[l2, y] * qE[l3, z] for x, y, z in perms) / 6
            assert lib.plijk_EEE(r, l1, l2, l3) == pytest.approx(want, rel=1e-15)
        assert lib.plijk_TTE(0, 9, 14, 41) == 0.0          # E leg above E lmax
        assert lib.plijk_TEE(0, 45, 14, 21) == 0.0         # T leg above T lmax
        # an E setup whose mode count differs from the l-space setup's
        ctx.lspace_set_E(E[0], E[1], E[2], E[3], tte[:-1], tee[:-1], eee[:-1])
        assert lib.calculate_orthol(0, 0, res) == -1
    finally:
        lib.mcmb_legacy_bind(None)


# ---- polarised cross-Gamma (calculate_gamma / calculate_gamma_2d result[0..3]) ----

def cross_case(Lq=40, T=(2, 40), E=(2, 36), RT=4, RE=4, npb=4, nx=26, seed=0):
    """non-separable q-tilde of T and E, TT/TE/EE spectra with s1 s4 - s3^2 > 0;
    the first 6 x samples scaled by 1e-12 (kept by the x < 10 rule, far below
    its 1e-18 threshold), samples 6..9 far above it (dropped)"""
    rng = np.random.default_rng(seed)
    ls = np.arange(Lq + 1.0)
    xv = np.sort(rng.uniform(0.2, 3.0, nx))
    qt = np.zeros((2, npb, Lq + 1, nx))
    for f in range(2):
        A = rng.standard_normal((npb, 3))
        for p in range(npb):
            for a in range(3):
                qt[f, p] += A[p, a] * np.outer(np.cos((a + 1 + f) * ls / Lq * (p + 1)), np.sin((a + 1) * xv + p + f))
    qt[:, :, :, :6] *= 1e-12
    ctt = np.where(ls > 0, 1000.0 / np.maximum(ls * (ls + 1), 1.0), 0.0)
    cee = 0.05 * ctt * (1 + 0.2 * np.cos(ls / 9))
    cte = 0.4 * np.sqrt(ctt * cee) * np.sin(ls / 5)
    spec = np.array([ctt, cte, cee, 0.02 + 0 * ls, 0.003 * np.cos(ls / 4), 0.03 + 0 * ls,
                     np.exp(-ls * (ls + 1) * 2e-4), np.exp(-ls * (ls + 1) * 3e-4), np.exp(-ls * (ls + 1) * 4e-4)])
    Tf, Ef = field(*T, RT, seed + 1), field(*E, RE, seed + 2)
    return dict(qtT=qt[0], qtE=qt[1], xv=xv, spec=spec, T=Tf, E=Ef, RT=RT, RE=RE, npb=npb, fr=0.8, dphi=1.5714e-8,
                prim=W.order_lex(npb)[rng.permutation(len(W.order_lex(npb)))].astype(np.int32))


def cross_setup(ctx, c):
    T, E = c["T"], c["E"]
    ttt, tte, tee, eee = setup(ctx, T, E, c["RT"], c["RE"], n=None)
    s = c["spec"]
    ctx.qtilde_set(c["qtT"], c["xv"], c["prim"], s[0], s[3], s[6], c["dphi"])
    ctx.qtilde_set_E(c["qtE"], s[1], s[2], s[4], s[5], s[7], s[8], c["fr"])
    return ttt, tte, tee, eee


@pytest.mark.parametrize("kw", [{}, dict(T=(5, 33), E=(2, 40), RT=5, RE=3, npb=3, seed=3)])
@pytest.mark.parametrize("spline", [modalcmb.SPLINE_FMM, modalcmb.SPLINE_NATURAL])
def test_cross_gl_pol_vs_oracle(ctx, kw, spline):
    c = cross_case(**kw)
    tabs = cross_setup(ctx, c)
    for pol in range(4):
        G = ctx.gamma_cross_gl_pol(pol, spline)
        ref = O.gamma_cross_gl_pol(pol, c["T"][:3], c["E"][:3], tabs[pol], c["qtT"], c["qtE"], c["xv"], c["prim"],
                                   c["spec"], c["fr"], c["dphi"], spline=spline)
        assert G.shape == ref.shape
        assert relmax(G, ref) < 1e-12, pol
    # TTT on the T lmax's node count is the TT cross-Gamma, bit for bit
    lt = c["T"][1]
    assert np.array_equal(ctx.gamma_cross_gl_pol(0, spline, 3 * lt // 2 + 1), ctx.gamma_cross_gl(spline))


def test_legacy_calculate_gamma_pol(ctx):
    lib = modalcmb.load_library()
    c = cross_case(RT=4, RE=4)
    RT, RE = 4, 4
    T, E = c["T"], c["E"]
    ttt = W.order_lex(RT).astype(np.int32)
    tte, tee, eee = tables(RT, RE, n=len(ttt))
    ctx.lspace_set(T[0], T[1], T[2], T[3], ttt, h2_mode=modalcmb.H2_ASYMPTOTIC)
    s = c["spec"]
    ctx.qtilde_set(c["qtT"], c["xv"], c["prim"], s[0], s[3], s[6], c["dphi"])
    lib.mcmb_legacy_bind(ctx._h)
    try:
        res = (C.c_double * 4)(-1, -1, -1, -1)
        lib.calculate_gamma_2d(2, 3, res)                  # T only: result[0]
        assert lib.mcmb_legacy_status() == 0 and res[1] == -1
        lib.precompute_gammaMij_E()
        assert lib.mcmb_legacy_status() == -7
        ctx.lspace_set_E(E[0], E[1], E[2], E[3], tte, tee, eee)
        ctx.qtilde_set_E(c["qtE"], s[1], s[2], s[4], s[5], s[7], s[8], c["fr"])
        lib.precompute_gammaMij_E()
        assert lib.mcmb_legacy_status() == 0
        lib.calculate_gammaMij_E(0, 0)
        assert lib.mcmb_legacy_status() == 0
        for spline, fn in ((modalcmb.SPLINE_FMM, lib.calculate_gamma_2d), (modalcmb.SPLINE_NATURAL, lib.calculate_gamma)):
            G = [ctx.gamma_cross_gl_pol(p, spline) for p in range(4)]
            for m, k in ((0, 0), (2, 3), (len(ttt) - 1, 5)):
                fn(m, k, res)
                assert lib.mcmb_legacy_status() == 0
                assert [res[q] for q in range(4)] == [G[q][m, k] for q in range(4)]
        # a new T q-tilde drops the E part: result[0] only again
        ctx.qtilde_set(c["qtT"] * 2.0, c["xv"], c["prim"], s[0], s[3], s[6], c["dphi"])
        res2 = (C.c_double * 4)(-1, -1, -1, -1)
        lib.calculate_gamma_2d(2, 3, res2)
        assert lib.mcmb_legacy_status() == 0 and res2[1] == -1
    finally:
        lib.mcmb_legacy_bind(None)


def test_cross_gl_pol_needs_e_setups(ctx):
    c = cross_case()
    T = c["T"]
    s = c["spec"]
    ctx.lspace_set(T[0], T[1], T[2], T[3], W.order_lex(c["RT"]), h2_mode=modalcmb.H2_ASYMPTOTIC)
    ctx.qtilde_set(c["qtT"], c["xv"], c["prim"], s[0], s[3], s[6], c["dphi"])
    ctx.nE = 1
    with pytest.raises(modalcmb.ModalError) as e:
        ctx.gamma_cross_gl_pol(modalcmb.POL_TEE)
    assert e.value.status == -7                            # no E late field
    E = c["E"]
    tte, tee, eee = tables(c["RT"], c["RE"])
    ctx.lspace_set_E(E[0], E[1], E[2], E[3], tte, tee, eee)
    with pytest.raises(modalcmb.ModalError) as e:
        ctx.gamma_cross_gl_pol(modalcmb.POL_TEE)
    assert e.value.status == -7                            # no E q-tilde
