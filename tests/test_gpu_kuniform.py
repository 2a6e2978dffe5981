"""GPU: the uniform k-grid tetrapyd of calculate_ortho / calculate_eigen
(reference parity mode) against the oracle's restatement: device point
weights bit-exact in walk order, the FP64 Gram and the projection to FP64
rounding, shares that sum to the whole, and the reference-named shims."""
import ctypes as C

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu

KMAX = 0.2


def setup(ctx, mx, R, mn=0):
    q = W.legendre_table(np.arange(mx + 1) / mx, R - 1)       # create_basis_prim: P_r(2 k/kmax - 1)
    order = W.order_lex(R)
    ctx.kuniform_set(mn, mx, KMAX, q, order)
    return q, order


def unpack(packed, n):
    g = np.zeros((n, n))
    iu = np.triu_indices(n)
    g[iu] = packed
    return g + np.triu(g, 1).T


@pytest.mark.parametrize("mx", [6, 31, 60])
def test_weights_bit_exact_in_walk_order(ctx, mx):
    setup(ctx, mx, 3)
    w = ctx.kuniform_weights()
    pts = O.walk_prim(0, mx)
    assert len(w) == len(pts) == ctx.kuniform_npoints()
    f = KMAX * (1.0 / mx)
    cube = f * f * f
    ref = np.empty(len(pts))
    for p, (i, j, k) in enumerate(pts.tolist()):
        x = O.lib().or_calculate_weight(0, mx, i, j, k) * cube
        ks = f * i + f * j + f * k
        if ks > 1e-10:
            ks = 1.0 / ks
        ref[p] = x * ks
    assert np.array_equal(w, ref)


def test_weights_with_nonzero_min(ctx):
    setup(ctx, 25, 3, mn=4)
    w = ctx.kuniform_weights()
    pts = O.walk_prim(4, 25)
    assert len(w) == len(pts)
    f = KMAX / 25
    for p in (0, 1, len(pts) // 2, len(pts) - 1):
        i, j, k = pts[p].tolist()
        x = O.lib().or_calculate_weight(4, 25, i, j, k) * (f * f * f) / (f * i + f * j + f * k)
        assert w[p] == pytest.approx(x, rel=1e-15)


@pytest.mark.parametrize("mx,R", [(40, 4), (75, 6)])
def test_gram_vs_oracle(ctx, mx, R):
    q, order = setup(ctx, mx, R)
    n = order.shape[0]
    G = ctx.kuniform_gram(modalcmb.FP64)
    packed = np.zeros(n * (n + 1) // 2)
    O.lib().or_calculate_ortho_uniform(mx, KMAX, O._p(np.ascontiguousarray(q)), q.shape[1], n,
                                       O._p(np.ascontiguousarray(order, dtype=np.int32)), O._p(packed))
    ref = unpack(packed, n)
    d = np.sqrt(np.outer(np.diag(ref), np.diag(ref)))
    assert np.array_equal(G, G.T)
    assert (np.abs(G - ref) / d).max() < 1e-12
    Gm = ctx.kuniform_gram(modalcmb.MIXED)
    assert (np.abs(Gm - ref) / d).max() < 2e-5


def test_gram_shares_sum_to_whole(ctx):
    setup(ctx, 50, 5)
    N = ctx.kuniform_npoints()
    G = ctx.kuniform_gram(modalcmb.FP64)
    cuts = [0, 1, 777, N // 2, N - 3, N]
    tot = sum(ctx.kuniform_gram(modalcmb.FP64, a, b) for a, b in zip(cuts[:-1], cuts[1:]))
    d = np.sqrt(np.outer(np.diag(G), np.diag(G)))
    assert (np.abs(tot - G) / d).max() < 1e-13
    with pytest.raises(modalcmb.ModalError):
        ctx.kuniform_gram(modalcmb.FP64, 5, N + 1)


def test_projection_vs_oracle(ctx):
    mx, R = 60, 5
    q, order = setup(ctx, mx, R)
    n = order.shape[0]
    params = [(150.0, 11.1), (37.5, 80.0), (410.0, 0.0), (12.0, 55.0), (250.0, 3.0), (90.0, 99.0), (5.0, 1.0),
              (333.0, 42.0), (64.0, 23.0)]        # (a1 = omega, a2 = phase in percent of 2 pi); 9 > one batch
    theta = np.array([(a1, 2.0 * 3.1415927 * a2 * 1e-2) for a1, a2 in params])
    A = ctx.kuniform_project(theta)
    for b, (a1, a2) in enumerate(params):
        ref = np.zeros(n)
        O.lib().or_calculate_eigen_uniform(mx, KMAX, O._p(np.ascontiguousarray(q)), q.shape[1], n,
                                           O._p(np.ascontiguousarray(order, dtype=np.int32)), C.c_double(a1),
                                           C.c_double(a2), O._p(ref))
        assert np.abs(A[b] - ref).max() < 1e-12 * np.abs(ref).max(), b
    # shares add up
    N = ctx.kuniform_npoints()
    part = ctx.kuniform_project(theta[:2], 0, 1000) + ctx.kuniform_project(theta[:2], 1000, N)
    assert np.abs(part - A[:2]).max() < 1e-13 * np.abs(A[:2]).max()


class ShapeParams(C.Structure):
    _fields_ = [("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double)]


def test_reference_named_shims_over_ranks_and_threads(ctx):
    """decompose_tetrapyd_prim for 3 ranks x 4 threads, then calculate_ortho
    and calculate_eigen per rank (accumulating, packed triangle / alpha):
    the ranks' sums equal the whole Gram and projection."""
    mx, R = 45, 4
    q, order = setup(ctx, mx, R)
    n = order.shape[0]
    lib = modalcmb.load_library()
    lib.calculate_eigen.argtypes = [C.c_int, C.c_int, ShapeParams, C.c_void_p, C.c_void_p]
    lib.calculate_ortho.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_void_p]
    lib.mcmb_legacy_bind(ctx._h)
    try:
        packed = np.zeros(n * (n + 1) // 2)
        alpha = np.zeros(n)
        sp = ShapeParams(150.0, 11.1, 0.0)
        for rank in range(3):
            blk = np.zeros((4, 4), dtype=np.int32)
            lib.decompose_tetrapyd_prim(blk.ctypes.data_as(C.c_void_p), rank, 3, 4, 0, mx)
            lib.calculate_ortho(0, mx, blk.ctypes.data_as(C.c_void_p), packed.ctypes.data_as(C.c_void_p))
            assert lib.mcmb_legacy_status() == 0
            lib.calculate_eigen(0, mx, sp, blk.ctypes.data_as(C.c_void_p), alpha.ctypes.data_as(C.c_void_p))
            assert lib.mcmb_legacy_status() == 0
        G = ctx.kuniform_gram(modalcmb.FP64)
        d = np.sqrt(np.outer(np.diag(G), np.diag(G)))
        assert (np.abs(unpack(packed, n) - G) / d).max() < 1e-13
        A = ctx.kuniform_project(np.array([[150.0, 2.0 * 3.1415927 * 11.1 * 1e-2]]))[0]
        assert np.abs(alpha - A).max() < 1e-13 * np.abs(A).max()
        # a mismatched grid is refused, not silently integrated
        blk = np.zeros((1, 4), dtype=np.int32)
        lib.decompose_tetrapyd_prim(blk.ctypes.data_as(C.c_void_p), 0, 1, 1, 0, mx + 1)
        lib.calculate_ortho(0, mx + 1, blk.ctypes.data_as(C.c_void_p), packed.ctypes.data_as(C.c_void_p))
        assert lib.mcmb_legacy_status() != 0
    finally:
        lib.mcmb_legacy_bind(None)
