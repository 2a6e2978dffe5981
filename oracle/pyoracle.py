"""ctypes wrapper of oracle/liboracle.so -- the FP64 CPU restatement of the
reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke test of the
repository's entry module and the CPU-baseline leg of bench.py, as the oracle
results are compared with.  Nothing in the product (modal-cmb_amd/) imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB):
        subprocess.run(["make", "-C", HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        L.or_count_tetrapyd_XXX.restype = C.c_long
        L.or_walk_XXX.restype = C.c_long
        L.or_walk_XXX.argtypes = [ip, ip, vp, C.c_long]
        L.or_walk_XXX_window.restype = C.c_long
        L.or_walk_XXX_window.argtypes = [ip, ip, C.c_long, vp, C.c_long]
        L.or_project_kgl_batch.argtypes = [ip, vp, vp, vp, ip, ip, vp, ip, vp, vp, vp, vp]
        L.or_qtilde3d.argtypes = [ip, ip, ip, vp, vp, vp, vp, vp, vp]
        L.or_gamma_cross_3d.argtypes = [ip, ip, ip, vp, ip, vp, ip, ip, vp, ip, vp, vp, ip, vp, ip, dp, C.c_long,
                                        C.c_long, vp]
        L.or_spline_integral_fmm.restype = dp
        L.or_spline_integral_fmm.argtypes = [vp, vp, ip]
        L.or_spline_integral_natural.restype = dp
        L.or_spline_integral_natural.argtypes = [vp, vp, ip]
        L.or_gamma_cross_gl.argtypes = [ip, ip, vp, ip, vp, ip, ip, ip, vp, ip, vp, vp, ip, vp, vp, vp, dp, ip, ip,
                                        vp]
        L.or_walk_prim.restype = C.c_long
        L.or_walk_prim.argtypes = [ip, ip, vp, C.c_long]
        for f in ("or_permsix", "or_geometric_asym", "or_threej000_sq", "or_h2_exact"):
            getattr(L, f).restype = dp
            getattr(L, f).argtypes = [ip, ip, ip]
        L.or_calculate_weight.restype = dp
        L.or_calculate_weight.argtypes = [ip, ip, ip, ip, ip]
        L.or_pijk.restype = dp
        L.or_feature.restype = dp
        L.or_feature.argtypes = [dp] * 5
        L.or_shape3_feature.restype = dp
        L.or_shape3_feature.argtypes = [dp] * 5
        L.or_wtime.restype = dp
        L.or_gamma_tet.argtypes = [ip, ip, vp, ip, vp, ip, vp, ip, ip, ip, vp]
        L.or_gamma_gl.argtypes = [ip, ip, vp, ip, vp, ip, vp, ip, vp]
        L.or_orthonormalise.argtypes = [ip, vp, vp, vp, vp]
        L.or_cholesky_solve.argtypes = [ip, vp, vp, vp]
        L.or_eigen_sym.argtypes = [ip, vp, vp, vp]
        L.or_ortho_late.argtypes = [ip, vp, ip, vp, vp, vp, vp]
        L.or_gram_kgl.argtypes = [ip, vp, vp, vp, ip, ip, vp, vp]
        L.or_project_kgl.argtypes = [ip, vp, vp, vp, ip, ip, vp, dp, dp, vp, vp, vp]
        L.or_calculate_ortho_uniform.argtypes = [ip, dp, vp, ip, ip, vp, vp]
        L.or_calculate_eigen_uniform.argtypes = [ip, dp, vp, ip, ip, vp, dp, dp, vp]
        L.or_basis_legendre.argtypes = [vp, ip, ip, dp, dp, vp]
        L.or_gauss_legendre.argtypes = [ip, vp, vp]
        L.or_threej_ftable.argtypes = [ip, vp]
        L.or_get_next_ijk_bound.argtypes = [ip, ip, ip, ip] + [C.POINTER(ip)] * 3 + [ip, ip, ip, C.POINTER(ip), ip]
        L.or_decompose_gamma_TTT_3D.argtypes = [ip, ip, ip, dp, ip, vp]
        L.or_decompose_tetrapyd_XXX.argtypes = [vp, ip, ip, ip, ip, ip]
        L.or_decompose_tetrapyd_prim.argtypes = [vp, ip, ip, ip, ip, ip]
        L.or_kgl_limits.argtypes = [ip, vp, vp]
        L.or_count_tetrapyd_XXY.restype = C.c_long
        L.or_count_tetrapyd_XXY.argtypes = [ip, ip, ip, ip]
        L.or_walk_XXY.restype = C.c_long
        L.or_walk_XXY.argtypes = [ip, ip, ip, ip, vp, C.c_long]
        L.or_decompose_tetrapyd_XXY.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
        L.or_get_ijk_next_XXY.argtypes = [ip, ip, ip, ip] + [C.POINTER(ip)] * 3
        L.or_check_ijk_min_XXY.argtypes = [ip, ip] + [C.POINTER(ip)] * 3
        L.or_check_ijk_max_XXY.argtypes = [ip, ip] + [C.POINTER(ip)] * 3
        L.or_gamma_gl_pol.argtypes = [ip, ip, ip, vp, ip, vp, ip, ip, vp, ip, vp, ip, vp, ip, vp]
        L.or_gamma_xxy_direct.argtypes = [ip, ip, ip, vp, ip, vp, ip, ip, vp, ip, vp, ip, vp, vp]
        L.or_gamma_cross_gl_pol.argtypes = [ip, ip, ip, vp, ip, ip, ip, vp, ip, vp, ip, ip, ip, vp, vp, ip, vp, vp,
                                            ip, vp, dp, dp, ip, ip, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def walk_XXX(lmin: int, lmax: int) -> np.ndarray:
    n = lib().or_count_tetrapyd_XXX(lmin, lmax)
    out = np.empty((n, 3), dtype=np.int32)
    lib().or_walk_XXX(lmin, lmax, _p(out), n)
    return out


def walk_XXX_window(lmin: int, lmax: int, p0: int, cap: int) -> np.ndarray:
    out = np.empty((cap, 3), dtype=np.int32)
    w = lib().or_walk_XXX_window(lmin, lmax, p0, _p(out), cap)
    return out[:w].copy()


def walk_prim(mn: int, mx: int) -> np.ndarray:
    cap = (mx + 1) ** 3
    buf = np.empty((cap, 3), dtype=np.int32)
    n = lib().or_walk_prim(mn, mx, _p(buf), cap)
    return buf[:n].copy()


def decompose_XXX(rank, nranks, nthreads, lmin, lmax) -> np.ndarray:
    out = np.zeros((nthreads, 4), dtype=np.int32)
    lib().or_decompose_tetrapyd_XXX(_p(out), rank, nranks, nthreads, lmin, lmax)
    return out


def decompose_prim(rank, nranks, nthreads, mn, mx) -> np.ndarray:
    out = np.zeros((nthreads, 4), dtype=np.int32)
    lib().or_decompose_tetrapyd_prim(_p(out), rank, nranks, nthreads, mn, mx)
    return out


def next_ijk_bound(lsize, i0, j0, k0, fi, fj, fk, maxblock):
    i1, j1, k1, bs = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    lib().or_get_next_ijk_bound(lsize, i0, j0, k0, C.byref(i1), C.byref(j1), C.byref(k1),
                                fi, fj, fk, C.byref(bs), maxblock)
    return i1.value, j1.value, k1.value, bs.value


def decompose_gamma_TTT_3D(lsize, rank, numranks, workshare, nthreads) -> np.ndarray:
    out = np.zeros((nthreads, 6), dtype=np.int32)
    lib().or_decompose_gamma_TTT_3D(lsize, rank, numranks, workshare, nthreads, _p(out))
    return out


def walk_XXY(lminX: int, lminY: int, lmaxX: int, lmaxY: int) -> np.ndarray:
    n = lib().or_count_tetrapyd_XXY(lminX, lminY, lmaxX, lmaxY)
    out = np.empty((n, 3), dtype=np.int32)
    lib().or_walk_XXY(lminX, lminY, lmaxX, lmaxY, _p(out), n)
    return out


def decompose_XXY(rank, nranks, nthreads, lminX, lminY, lmaxX, lmaxY) -> np.ndarray:
    out = np.zeros((nthreads, 4), dtype=np.int32)
    lib().or_decompose_tetrapyd_XXY(_p(out), rank, nranks, nthreads, lminX, lminY, lmaxX, lmaxY)
    return out


def _ijk_call(f, a, b, ijk):
    i, j, k = (C.c_int(int(x)) for x in ijk)
    r = f(a, b, C.byref(i), C.byref(j), C.byref(k))
    return r, (i.value, j.value, k.value)


def next_XXY(lminX, lminY, lmaxX, lmaxY, ijk):
    """(more, (i, j, k)) after one get_ijk_next_XXY step."""
    i, j, k = (C.c_int(int(x)) for x in ijk)
    r = lib().or_get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, C.byref(i), C.byref(j), C.byref(k))
    return r, (i.value, j.value, k.value)


def check_min_XXY(lminX, lminY, ijk):
    return _ijk_call(lib().or_check_ijk_min_XXY, lminX, lminY, ijk)[1]


def check_max_XXY(lmaxX, lmaxY, ijk):
    return _ijk_call(lib().or_check_ijk_max_XXY, lmaxX, lmaxY, ijk)[1]


def _pol_args(T, E):
    """(lmin, lmax, q, v) of each field -> flattened ctypes arguments"""
    out = []
    for lmin, lmax, q, v in (T, E):
        q = np.ascontiguousarray(q, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        out.append((lmin, lmax, q, v))
    return out


def gamma_gl_pol(pol, T, E, order, npt=None) -> np.ndarray:
    """calculate_orthol's component pol (0 TTT, 1 TTE, 2 TEE, 3 EEE); T and E
    are (lmin, lmax, q[l][R], v[l]) of each field."""
    (tl, th, qT, vT), (el, eh, qE, vE) = _pol_args(T, E)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    npt = 3 * max(th, eh) // 2 + 1 if npt is None else npt
    g = np.empty((n, n))
    lib().or_gamma_gl_pol(pol, tl, th, _p(qT), qT.shape[1], _p(vT), el, eh, _p(qE), qE.shape[1], _p(vE), n,
                          _p(order), npt, _p(g))
    return g


def gamma_xxy_direct(pol, T, E, order) -> np.ndarray:
    (tl, th, qT, vT), (el, eh, qE, vE) = _pol_args(T, E)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    g = np.empty((n, n))
    lib().or_gamma_xxy_direct(pol, tl, th, _p(qT), qT.shape[1], _p(vT), el, eh, _p(qE), qE.shape[1], _p(vE), n,
                              _p(order), _p(g))
    return g


def gamma_cross_gl_pol(pol, T, E, order_late, qtT, qtE, xvec, order_prim, spec, fsky_ratio, dphi, spline=0,
                       npt=None) -> np.ndarray:
    """calculate_gamma(_2d)'s component pol; T / E = (lmin, lmax, b[l][R]) of
    the late fields, qtT / qtE [np][lsize_q][nx], spec [9][lsize_q] (C_TT, C_TE,
    C_EE, N_TT, N_TE, N_EE, b_TT, b_TE, b_EE)."""
    (tl, th, bT), (el, eh, bE) = T, E
    bT = np.ascontiguousarray(bT, dtype=np.float64)
    bE = np.ascontiguousarray(bE, dtype=np.float64)
    ol = np.ascontiguousarray(order_late, dtype=np.int32)
    op = np.ascontiguousarray(order_prim, dtype=np.int32)
    qtT = np.ascontiguousarray(qtT, dtype=np.float64)
    qtE = np.ascontiguousarray(qtE, dtype=np.float64)
    xvec = np.ascontiguousarray(xvec, dtype=np.float64)
    spec = np.ascontiguousarray(spec, dtype=np.float64)
    np_, lsq, nx = qtT.shape
    assert qtE.shape == qtT.shape and spec.shape == (9, lsq)
    npt = 3 * max(th, eh) // 2 + 1 if npt is None else npt
    g = np.empty((ol.shape[0], op.shape[0]))
    lib().or_gamma_cross_gl_pol(pol, tl, th, _p(bT), bT.shape[1], el, eh, _p(bE), bE.shape[1], _p(ol), ol.shape[0],
                                np_, lsq, _p(qtT), _p(qtE), nx, _p(xvec), _p(op), op.shape[0], _p(spec),
                                float(fsky_ratio), float(dphi), spline, npt, _p(g))
    return g


def basis_legendre(values: np.ndarray, pmax: int, vmin: float, vmax: float) -> np.ndarray:
    values = np.ascontiguousarray(values, dtype=np.float64)
    out = np.empty((values.shape[0], pmax + 1))
    lib().or_basis_legendre(_p(out), values.shape[0], pmax, vmin, vmax, _p(values))
    return out


def gauss_legendre(n: int):
    x, w = np.empty(n), np.empty(n)
    lib().or_gauss_legendre(n, _p(x), _p(w))
    return x, w


def ftable(amax: int) -> np.ndarray:
    f = np.empty(amax + 1)
    lib().or_threej_ftable(amax, _p(f))
    return f


def gamma_tet(lmin, lmax, q, v, order, geom=0, i_lo=0, i_hi=0) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    g = np.empty((n, n))
    lib().or_gamma_tet(lmin, lmax, _p(q), q.shape[1], _p(v), n, _p(order), geom, i_lo, i_hi, _p(g))
    return g


def gamma_gl(lmin, lmax, q, v, order, npt=None) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    npt = 3 * lmax // 2 + 1 if npt is None else npt
    g = np.empty((n, n))
    lib().or_gamma_gl(lmin, lmax, _p(q), q.shape[1], _p(v), n, _p(order), npt, _p(g))
    return g


def orthonormalise(gamma: np.ndarray):
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    n = g.shape[0]
    ortho, lam, norm = np.empty_like(g), np.empty_like(g), np.empty(n)
    st = lib().or_orthonormalise(n, _p(g), _p(ortho), _p(lam), _p(norm))
    return st, ortho, lam, norm


def eigen_sym(a: np.ndarray):
    """(sweeps, evals descending, evecs[:, k]) of a symmetric matrix."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    e, v = np.empty(n), np.empty_like(a)
    st = lib().or_eigen_sym(n, _p(a), _p(e), _p(v))
    return st, e, v


def ortho_late(gamma: np.ndarray, pca_flag: int = 1):
    """(r, lambda, lambdainv, norm, evals) as ortho_late_mpi.c:752-842 leaves them."""
    g = np.ascontiguousarray(gamma, dtype=np.float64)
    n = g.shape[0]
    lam, linv, norm, ev = np.empty_like(g), np.empty_like(g), np.empty(n), np.empty(n)
    r = lib().or_ortho_late(n, _p(g), pca_flag, _p(lam), _p(linv), _p(norm), _p(ev))
    return r, lam, linv, norm, ev


def cholesky_solve(ortho: np.ndarray, alpha: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(ortho, dtype=np.float64)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.empty_like(a)
    lib().or_cholesky_solve(f.shape[0], _p(f), _p(a), _p(b))
    return b


def gram_kgl(kn, kw, q, order) -> np.ndarray:
    kn = np.ascontiguousarray(kn, dtype=np.float64)
    kw = np.ascontiguousarray(kw, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    g = np.empty((n, n))
    lib().or_gram_kgl(kn.shape[0], _p(kn), _p(kw), _p(q), q.shape[1], n, _p(order), _p(g))
    return g


def project_kgl(kn, kw, q, order, omega, phi):
    kn = np.ascontiguousarray(kn, dtype=np.float64)
    kw = np.ascontiguousarray(kw, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    a, d1, d2 = np.empty(n), np.empty(n), np.empty(n)
    lib().or_project_kgl(kn.shape[0], _p(kn), _p(kw), _p(q), q.shape[1], n, _p(order),
                         float(omega), float(phi), _p(a), _p(d1), _p(d2))
    return a, d1, d2


def project_kgl_batch(kn, kw, q, order, theta):
    """alpha[b][n], d alpha / d omega [b][n], d alpha / d phi [b][n] for theta[b] = (omega, phi)."""
    kn = np.ascontiguousarray(kn, dtype=np.float64)
    kw = np.ascontiguousarray(kw, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1, 2)
    n, B = order.shape[0], th.shape[0]
    a, d1, d2 = np.empty((B, n)), np.empty((B, n)), np.empty((B, n))
    lib().or_project_kgl_batch(kn.shape[0], _p(kn), _p(kw), _p(q), q.shape[1], n, _p(order), B, _p(th),
                               _p(a), _p(d1), _p(d2))
    return a, d1, d2


def kgl_limits(kn) -> np.ndarray:
    kn = np.ascontiguousarray(kn, dtype=np.float64)
    N = kn.shape[0]
    lim = np.empty((N, N), dtype=np.int32)
    lib().or_kgl_limits(N, _p(kn), _p(lim))
    return lim


def calculate_ortho_uniform(mx, kmax, q, order) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    out = np.zeros(n * (n + 1) // 2)
    lib().or_calculate_ortho_uniform(mx, float(kmax), _p(q), q.shape[1], n, _p(order), _p(out))
    return out


def calculate_eigen_uniform(mx, kmax, q, order, a1, a2) -> np.ndarray:
    q = np.ascontiguousarray(q, dtype=np.float64)
    order = np.ascontiguousarray(order, dtype=np.int32)
    n = order.shape[0]
    out = np.zeros(n)
    lib().or_calculate_eigen_uniform(mx, float(kmax), _p(q), q.shape[1], n, _p(order),
                                     float(a1), float(a2), _p(out))
    return out


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def qtilde3d(qt, cl, nl, bl):
    """(q3[l][p][x], lw[l]) of init_gamma_3Dint from qt[p][l][x] and the spectra."""
    qt = _c(qt)
    np_, lsize, nx = qt.shape
    q3, lw = np.empty((lsize, np_, nx)), np.empty(lsize)
    lib().or_qtilde3d(np_, lsize, nx, _p(qt), _p(_c(cl)), _p(_c(nl)), _p(_c(bl)), _p(q3), _p(lw))
    return q3, lw


def gamma_cross_3d(tlmin, tlmax, b, order_late, q3, xvec, order_prim, lw, dphi, geom=1, p0=0, p1=0):
    b, q3, xvec, lw = _c(b), _c(q3), _c(xvec), _c(lw)
    ol, op = _c(order_late, np.int32), _c(order_prim, np.int32)
    lsize, np_, nx = q3.shape
    nl_, n = ol.shape[0], op.shape[0]
    g = np.empty((nl_, n))
    lib().or_gamma_cross_3d(lsize, tlmin, tlmax, _p(b), b.shape[1], _p(ol), nl_, np_, _p(q3), nx, _p(xvec), _p(op),
                            n, _p(lw), geom, float(dphi), int(p0), int(p1), _p(g))
    return g


def spline_integral(x, y, kind="fmm"):
    x, y = _c(x), _c(y)
    f = lib().or_spline_integral_fmm if kind == "fmm" else lib().or_spline_integral_natural
    return f(_p(x), _p(y), x.shape[0])


def gamma_cross_gl(tlmin, tlmax, b, order_late, qt, xvec, order_prim, cl, nl, bl, dphi, spline=0, npt=None):
    b, qt, xvec = _c(b), _c(qt), _c(xvec)
    ol, op = _c(order_late, np.int32), _c(order_prim, np.int32)
    np_, lsize_q, nx = qt.shape
    nl_, n = ol.shape[0], op.shape[0]
    npt = 3 * tlmax // 2 + 1 if npt is None else npt
    g = np.empty((nl_, n))
    lib().or_gamma_cross_gl(tlmin, tlmax, _p(b), b.shape[1], _p(ol), nl_, np_, lsize_q, _p(qt), nx, _p(xvec),
                            _p(op), n, _p(_c(cl)), _p(_c(nl)), _p(_c(bl)), float(dphi), spline, npt, _p(g))
    return g
