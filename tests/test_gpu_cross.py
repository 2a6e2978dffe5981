"""GPU cross-Gamma (late basis against the projected primordial basis
q-tilde) through the C ABI against the oracle: the 3D form of gamma_3d_host
(trapezium in x) and the Gauss-Legendre form of calculate_gamma_2d (FMM
spline) and calculate_gamma (natural spline), with the reference-named
entry points over a bound context."""
import ctypes as C
import threading

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def relmax(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def case(Lq=40, tlmin=2, tlmax=40, R=4, npb=4, nx=30, seed=0):
    """A late Legendre basis on [tlmin, tlmax] and a q-tilde that does not
    factorise in x: qt[p][l][x] = sum of a few smooth products; the samples
    x < 5 are scaled by 1e-9 (|y| far below 1e-18: kept), 5 <= x < 10 are not
    (|y| above 1e-18: dropped)."""
    rng = np.random.default_rng(seed)
    ls = np.arange(Lq + 1.0)
    b = np.zeros((Lq + 1, R))
    b[tlmin:tlmax + 1] = W.legendre_table((ls[tlmin:tlmax + 1] - tlmin) / (tlmax - tlmin), R - 1)
    xvec = np.sort(rng.uniform(0.2, 3.0, nx))
    A = rng.standard_normal((npb, 3))
    qt = np.zeros((npb, Lq + 1, nx))
    for p in range(npb):
        for a in range(3):
            qt[p] += A[p, a] * np.outer(np.cos((a + 1) * ls / Lq * (p + 1)), np.sin((a + 1) * xvec + p))
    qt[:, :, :5] *= 1e-9
    cl = np.zeros(Lq + 1)
    cl[1:] = 1000.0 / (ls[1:] * (ls[1:] + 1.0))
    nl = 0.02 * np.ones(Lq + 1)
    bl = np.exp(-ls * (ls + 1) * 2e-4)
    order_late = W.order_lex(R)
    order_prim = W.order_lex(npb)[rng.permutation(len(W.order_lex(npb)))]
    return dict(Lq=Lq, tlmin=tlmin, tlmax=tlmax, b=b, qt=qt, xvec=xvec, cl=cl, nl=nl, bl=bl,
                order_late=order_late, order_prim=order_prim, dphi=1.5714e-8)


def setup(ctx, c, h2=modalcmb.H2_ASYMPTOTIC):
    v = np.ones(c["tlmax"] + 1)   # the auto-Gram weight; the cross forms do not read it
    ctx.lspace_set(c["tlmin"], c["tlmax"], c["b"][:c["tlmax"] + 1], v, c["order_late"], h2_mode=h2)
    ctx.qtilde_set(c["qt"], c["xvec"], c["order_prim"], c["cl"], c["nl"], c["bl"], c["dphi"])


def oracle_3d(c, geom, p0=0, p1=0):
    q3, lw = O.qtilde3d(c["qt"], c["cl"], c["nl"], c["bl"])
    return O.gamma_cross_3d(c["tlmin"], c["tlmax"], c["b"], c["order_late"], q3, c["xvec"], c["order_prim"], lw,
                            c["dphi"], geom=geom, p0=p0, p1=p1)


@pytest.mark.parametrize("kw,h2", [({}, modalcmb.H2_ASYMPTOTIC), ({}, modalcmb.H2_EXACT),
                                   (dict(tlmin=5, tlmax=33, R=5, npb=3), modalcmb.H2_ASYMPTOTIC),
                                   (dict(Lq=57, tlmax=57, R=6, npb=6, nx=41, seed=3), modalcmb.H2_ASYMPTOTIC)])
def test_cross_3d_vs_oracle(ctx, kw, h2):
    c = case(**kw)
    setup(ctx, c, h2)
    G = ctx.gamma_cross_3d()
    ref = oracle_3d(c, 1 if h2 == modalcmb.H2_ASYMPTOTIC else 0)
    assert G.shape == ref.shape
    assert relmax(G, ref) < 1e-12


def test_cross_3d_rank_shares(ctx):
    c = case()
    setup(ctx, c)
    G = ctx.gamma_cross_3d()
    N = len(O.walk_XXX(2, c["Lq"]))
    parts = [ctx.gamma_cross_3d(r, 3) for r in range(3)]
    assert relmax(sum(parts), G) < 1e-13
    # rank 1 of 3: points [N/3 + (1 < N%3), ...) of the walk, as
    # decompose_gamma_TTT_3D_mpi splits them
    blk, left = divmod(N, 3)
    p0 = blk + min(1, left)
    p1 = p0 + blk + (1 < left)
    assert relmax(parts[1], oracle_3d(c, 1, p0, p1)) < 1e-12


@pytest.mark.parametrize("spline", [modalcmb.SPLINE_FMM, modalcmb.SPLINE_NATURAL])
@pytest.mark.parametrize("kw", [{}, dict(tlmin=5, tlmax=33, R=5, npb=3), dict(Lq=64, tlmax=60, R=7, npb=5, nx=37)])
def test_cross_gl_vs_oracle(ctx, spline, kw):
    c = case(**kw)
    setup(ctx, c)
    G = ctx.gamma_cross_gl(spline)
    ref = O.gamma_cross_gl(c["tlmin"], c["tlmax"], c["b"], c["order_late"], c["qt"], c["xvec"], c["order_prim"],
                           c["cl"], c["nl"], c["bl"], c["dphi"], spline=spline)
    assert relmax(G, ref) < 1e-12


def test_legacy_cross_entry_points(ctx):
    lib = modalcmb.load_library()
    c = case()
    setup(ctx, c)
    n = len(c["order_late"])
    assert n == len(c["order_prim"])
    G3 = ctx.gamma_cross_3d()
    Gf = ctx.gamma_cross_gl(modalcmb.SPLINE_FMM)
    Gn = ctx.gamma_cross_gl(modalcmb.SPLINE_NATURAL)
    lib.mcmb_legacy_bind(ctx._h)
    try:
        buf = np.full((n, n), 7.0)
        lib.gamma_3d_host(buf.ctypes.data_as(C.c_void_p), n, 0, 1, C.c_double(0.0))
        assert lib.mcmb_legacy_status() == 0
        assert np.array_equal(buf, G3)
        tot = np.zeros((n, n))
        for r in range(4):
            lib.gamma_3d_host(buf.ctypes.data_as(C.c_void_p), n, r, 4, C.c_double(0.0))
            tot += buf
        assert relmax(tot, G3) < 1e-13
        assert lib.get_qtilde_xsize() == len(c["xvec"]) and lib.get_qtilde_T_lsize() == c["Lq"] + 1
        lib.init_gamma_glint()
        assert lib.mcmb_legacy_status() == 0
        # calculate_gamma_2d is called inside an omp for by gamma_mpi.c: many
        # threads at once, the first fills the cache
        out = np.zeros((n, n))
        bad = []

        def work(rows):
            res = (C.c_double * 4)()
            for m in rows:
                for k in range(n):
                    lib.calculate_gamma_2d(m, k, res)
                    if lib.mcmb_legacy_status() != 0:
                        bad.append((m, k))
                    out[m, k] = res[0]
        th = [threading.Thread(target=work, args=(range(t, n, 4),)) for t in range(4)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not bad
        assert np.array_equal(out, Gf)
        res = (C.c_double * 4)()
        lib.calculate_gamma(3, 5, res)
        assert lib.mcmb_legacy_status() == 0 and res[0] == Gn[3, 5]
        # a new q-tilde setup drops the cached Gamma
        c2 = dict(c, qt=c["qt"] * 2.0)
        ctx.qtilde_set(c2["qt"], c2["xvec"], c2["order_prim"], c2["cl"], c2["nl"], c2["bl"], c2["dphi"])
        lib.calculate_gamma_2d(3, 5, res)
        assert res[0] == pytest.approx(8.0 * Gf[3, 5], rel=1e-12)
    finally:
        lib.mcmb_legacy_bind(None)


def test_legacy_gamma_3d_host_needs_qtilde():
    lib = modalcmb.load_library()
    with modalcmb.Context(0) as c2:
        ws = W.lspace(30, 3)
        c2.lspace_set(2, 30, ws["q"], ws["v"], ws["order"])
        lib.mcmb_legacy_bind(c2._h)
        try:
            n = len(ws["order"])
            buf = np.full((n, n), 7.0)
            lib.gamma_3d_host(buf.ctypes.data_as(C.c_void_p), n, 0, 1, C.c_double(0.0))
            assert lib.mcmb_legacy_status() == -7    # MCMB_ERR_STATE
            assert not buf.any()
        finally:
            lib.mcmb_legacy_bind(None)
