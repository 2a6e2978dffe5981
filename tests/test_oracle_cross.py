"""Pins of the cross-Gamma oracle (late basis against the projected
primordial basis q-tilde; gamma_3d_host and calculate_gamma_2d /
calculate_gamma), CPU only.

* Spline rules: the FMM spline of integrate_cubic_spline is exact on cubics
  and equals a spline built here from its definition (knot second derivatives
  from the continuity equations, end rows matching the third divided
  differences of the first and last four points); the natural spline equals
  scipy's CubicSpline(bc_type="natural").
* Both cross forms against the pinned auto-Gram oracles: for q-tilde that
  factorises as qt[p][l][x] = b_{pi(p)}(l) g(x), the mode products factorise
  and Gamma[m][n] = 6 dphi^2 I[x^2 g^3] Gamma_auto[m][n'] with the per-leg
  weight lw(l) s4(l) = (2l+1)^(-1/6) b_l / sqrt(b_l^2 C_l + N_l), n' the late
  mode with the sorted indices pi(order_prim[n]), and I the trapezium rule
  (3D form) or the spline rule (GL forms).  g is chosen so that the samples
  x < 10 cover both sides of the 1e-18 rule: tiny ones are kept, large ones
  dropped.
"""
import numpy as np
import pytest

import pyoracle as O
from modalcmb import workloads as W

try:
    from scipy.interpolate import CubicSpline
except ImportError:  # pragma: no cover
    CubicSpline = None


def fmm_by_definition(x, y):
    n = len(x)
    h = np.diff(x)
    A = np.zeros((n, n))
    r = np.zeros(n)
    for i in range(1, n - 1):
        A[i, i - 1], A[i, i], A[i, i + 1] = h[i - 1], 2 * (h[i - 1] + h[i]), h[i]
        r[i] = 6 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1])

    def dd3(xs, ys):   # third divided difference
        d1 = np.diff(ys) / np.diff(xs)
        d2 = np.diff(d1) / (xs[2:] - xs[:-2])
        return (d2[1] - d2[0]) / (xs[3] - xs[0])
    # s''' on the end intervals = 6 f[x0..x3] and 6 f[x_{n-4}..x_{n-1}]
    A[0, 0], A[0, 1], r[0] = -1.0 / h[0], 1.0 / h[0], 6 * dd3(x[:4], y[:4])
    A[-1, -2], A[-1, -1], r[-1] = -1.0 / h[-1], 1.0 / h[-1], 6 * dd3(x[-4:], y[-4:])
    M = np.linalg.solve(A, r)
    return float(np.sum(h * (y[:-1] + y[1:]) / 2 - h ** 3 * (M[:-1] + M[1:]) / 24))


def test_spline_rules():
    rng = np.random.default_rng(2)
    x = np.sort(rng.uniform(0, 5, 23))
    for c in ([1.0, -2.0, 0.5, 0.25], [0.0, 0.0, 0.0, 1.0]):
        y = np.polyval(c[::-1], x)
        exact = np.polyval(np.polyint(c[::-1]), x[-1]) - np.polyval(np.polyint(c[::-1]), x[0])
        assert O.spline_integral(x, y, "fmm") == pytest.approx(exact, rel=1e-13)
    y = rng.standard_normal(23)
    assert O.spline_integral(x, y, "fmm") == pytest.approx(fmm_by_definition(x, y), rel=1e-12, abs=1e-12)
    if CubicSpline is not None:
        ref = CubicSpline(x, y, bc_type="natural").integrate(x[0], x[-1])
        assert O.spline_integral(x, y, "natural") == pytest.approx(ref, rel=1e-12, abs=1e-12)
    # two points: the natural spline is the trapezium (the FMM end conditions
    # need four points; with two the reference's elimination divides 0 by 0)
    assert O.spline_integral(x[:2], y[:2], "natural") == pytest.approx((x[1] - x[0]) * (y[0] + y[1]) / 2)
    assert np.isnan(O.spline_integral(x[:3], y[:3], "fmm"))


def _separable_case(L=30, R=4, nx=24, seed=0):
    rng = np.random.default_rng(seed)
    ls = np.arange(L + 1.0)
    b = W.legendre_table(ls / L, R - 1)                    # late basis, rows by multipole
    perm = rng.permutation(R)                              # q-tilde_p = b_{perm[p]} g
    xvec = np.sort(rng.uniform(0.0, 6.0, nx))
    g = np.ones(nx)
    g[:5] = 1e-9           # x < 10 with |y| far below 1e-18: kept
    g[5:10] = 1.0          # x < 10 with |y| above 1e-18: dropped
    g[10:] = 0.5 + 0.3 * np.sin(xvec[10:])
    qt = np.einsum("lp,x->plx", b[:, perm], g)
    cl = np.zeros(L + 1)
    cl[1:] = 1000.0 / (ls[1:] * (ls[1:] + 1.0))
    nl = 0.05 * np.ones(L + 1)
    bl = np.exp(-ls * (ls + 1) * 1e-4)
    order_late = W.order_lex(R)
    # primordial modes: the late table with its slots shuffled, in another order
    order_prim = order_late[rng.permutation(len(order_late))]
    order_prim = np.array([o[rng.permutation(3)] for o in order_prim], dtype=np.int32)
    # n' of each primordial mode in the late table
    idx = {tuple(o): i for i, o in enumerate(order_late.tolist())}
    nprime = np.array([idx[tuple(sorted(perm[o]))] for o in order_prim.tolist()])
    v = np.zeros(L + 1)
    v[2:] = (2 * ls[2:] + 1) ** (-1 / 6) * bl[2:] / np.sqrt(bl[2:] ** 2 * cl[2:] + nl[2:])
    y = xvec ** 2 * g ** 3
    ym = y.copy()
    ym[5:10] = 0.0
    return dict(L=L, b=b, qt=qt, xvec=xvec, cl=cl, nl=nl, bl=bl, order_late=order_late, order_prim=order_prim,
                nprime=nprime, v=v, ymask=ym)


@pytest.mark.parametrize("geom", [0, 1])
def test_cross_3d_factorised_qtilde(geom):
    c = _separable_case()
    L, dphi = c["L"], 1.5714e-8
    q3, lw = O.qtilde3d(c["qt"], c["cl"], c["nl"], c["bl"])
    assert np.allclose(lw[2:] * q3[2:, 0, 10] / c["qt"][0, 2:, 10],
                       c["v"][2:], rtol=1e-14, atol=0)
    G = O.gamma_cross_3d(2, L, c["b"], c["order_late"], q3, c["xvec"], c["order_prim"], lw, dphi, geom=geom)
    auto = O.gamma_tet(2, L, c["b"], c["v"], c["order_late"], geom=geom)
    x, ym = c["xvec"], c["ymask"]
    trap = float(np.sum(np.diff(x) * (ym[:-1] + ym[1:]) / 2))
    ref = 6 * dphi ** 2 * trap * auto[:, c["nprime"]]
    assert np.abs(G - ref).max() < 1e-12 * np.abs(ref).max()
    # rank shares of the walk (decompose_gamma_TTT_3D_mpi's split) sum to it
    N = len(O.walk_XXX(2, L))
    parts = [O.gamma_cross_3d(2, L, c["b"], c["order_late"], q3, c["xvec"], c["order_prim"], lw, dphi,
                              geom=geom, p0=a, p1=e) for a, e in [(0, N // 3), (N // 3, N)]]
    assert np.abs(parts[0] + parts[1] - G).max() < 1e-13 * np.abs(G).max()


@pytest.mark.parametrize("spline", [0, 1])
def test_cross_gl_factorised_qtilde(spline):
    c = _separable_case()
    L, dphi = c["L"], 1.5714e-8
    G = O.gamma_cross_gl(2, L, c["b"], c["order_late"], c["qt"], c["xvec"], c["order_prim"], c["cl"], c["nl"],
                         c["bl"], dphi, spline=spline)
    auto = O.gamma_gl(2, L, c["b"], c["v"], c["order_late"])
    S = O.spline_integral(c["xvec"], c["ymask"], "fmm" if spline == 0 else "natural")
    ref = 6 * dphi ** 2 * S * auto[:, c["nprime"]]
    assert np.abs(G - ref).max() < 1e-12 * np.abs(ref).max()
    # the exact-h2 3D form with the same rule in x is the same number
    q3, lw = O.qtilde3d(c["qt"], c["cl"], c["nl"], c["bl"])
    G3 = O.gamma_cross_3d(2, L, c["b"], c["order_late"], q3, c["xvec"], c["order_prim"], lw, dphi, geom=0)
    trap = float(np.sum(np.diff(c["xvec"]) * (c["ymask"][:-1] + c["ymask"][1:]) / 2))
    assert np.abs(G / S - G3 / trap).max() < 1e-11 * np.abs(G / S).max()


def test_cross_gl_late_range_and_general_qtilde():
    # a late range inside the q-tilde's l range, q-tilde not factorised:
    # the GL form equals the exact-h2 3D form when both integrate in x with
    # the same weights per sample (here: one sample at a time, x >= 10 only)
    rng = np.random.default_rng(4)
    L, R, nx = 26, 3, 14
    ls = np.arange(L + 1.0)
    b = W.legendre_table((ls - 4.0) / (L - 4.0), R - 1)
    qt = rng.standard_normal((R, L + 1, nx))
    cl = np.ones(L + 1)
    nl = np.zeros(L + 1)
    bl = np.ones(L + 1)
    order = W.order_lex(R)
    xvec = np.linspace(1.0, 2.0, nx)
    q3, lw = O.qtilde3d(qt, cl, nl, bl)
    dphi = 1.0
    for xi in (10, 13):
        # isolate sample xi: a q-tilde that is zero at every other x
        qx = np.zeros_like(qt)
        qx[:, :, xi] = qt[:, :, xi]
        q3x = np.zeros_like(q3)
        q3x[:, :, xi] = q3[:, :, xi]
        Gg = O.gamma_cross_gl(4, 22, b, order, qx, xvec, order, cl, nl, bl, dphi, spline=0)
        G3 = O.gamma_cross_3d(4, 22, b, order, q3x, xvec, order, lw, dphi, geom=0)
        wg = O.spline_integral(xvec, np.eye(nx)[xi], "fmm")
        wt = float(np.sum(np.diff(xvec) * (np.eye(nx)[xi][:-1] + np.eye(nx)[xi][1:]) / 2))
        assert np.abs(Gg / wg - G3 / wt).max() < 1e-11 * np.abs(G3 / wt).max()
