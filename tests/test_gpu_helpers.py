"""GPU: reference-named helpers over the bound context (pijk, plijk_TTT,
find_perm_prim, get_terms_prim, get_kmax, create_basis_prim) against the
oracle, and the Gamma post-processing against numpy."""
import ctypes as C

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def test_prim_helpers_match_oracle(ctx):
    mx, R = 30, 4
    q = W.legendre_table(np.arange(mx + 1) / mx, R - 1)
    order = W.order_lex(R)
    ctx.kuniform_set(0, mx, 0.2, q, order)
    lib = modalcmb.load_library()
    lib.mcmb_legacy_bind(ctx._h)
    try:
        assert lib.get_terms_prim() == order.shape[0]
        assert lib.get_kmax() == 0.2
        p = [C.c_int(), C.c_int(), C.c_int()]
        lib.find_perm_prim(5, *[C.byref(x) for x in p])
        assert [x.value for x in p] == order[5].tolist()
        qo, oo = np.ascontiguousarray(q), np.ascontiguousarray(order, dtype=np.int32)
        rng = np.random.default_rng(0)
        for _ in range(40):
            r = int(rng.integers(order.shape[0]))
            i, j, k = (int(v) for v in rng.integers(0, mx + 1, 3))
            ref = O.lib().or_pijk(O._p(qo), q.shape[1], O._p(oo), r, i, j, k)
            assert lib.pijk(r, i, j, k) == ref
        # create_basis_prim rebuilds the grid and basis, keeping the order table
        vec = 0.25 * np.arange(41) / 40
        assert lib.create_basis_prim(41, C.c_double(0.25), vec.ctypes.data_as(C.c_void_p)) == 0
        assert lib.get_kmax() == 0.25 and ctx.kuniform_npoints() == len(O.walk_prim(0, 40))
        q2 = W.legendre_table(vec / 0.25, R - 1)
        ref = O.lib().or_pijk(O._p(np.ascontiguousarray(q2)), q2.shape[1], O._p(oo), 3, 7, 9, 12)
        assert lib.pijk(3, 7, 9, 12) == pytest.approx(ref, rel=1e-14)
    finally:
        lib.mcmb_legacy_bind(None)


def test_plijk_ttt_matches_oracle_and_masks_range(ctx):
    L, R = 60, 4
    ws = W.lspace(L, R, True)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    lib = modalcmb.load_library()
    lib.mcmb_legacy_bind(ctx._h)
    try:
        qo, oo = np.ascontiguousarray(ws["q"]), np.ascontiguousarray(ws["order"], dtype=np.int32)
        for (r, a, b, c) in [(0, 2, 2, 2), (7, 10, 33, 40), (19, 60, 60, 60), (11, 5, 17, 21)]:
            assert lib.plijk_TTT(r, a, b, c) == O.lib().or_pijk(O._p(qo), ws["q"].shape[1], O._p(oo), r, a, b, c)
        assert lib.plijk_TTT(3, 1, 5, 5) == 0.0 and lib.plijk_TTT(3, 5, 5, 61) == 0.0
    finally:
        lib.mcmb_legacy_bind(None)


@pytest.mark.parametrize("n", [1, 17, 130, 600])
def test_gamma_post_against_numpy(ctx, n):
    rng = np.random.default_rng(n)
    lam = rng.standard_normal((n, n))
    G = rng.standard_normal((n, n))
    G = G @ G.T
    oi, out = ctx.gamma_post(lam, G)
    ref_oi = lam.T @ lam
    assert np.abs(oi - ref_oi).max() < 1e-12 * np.abs(ref_oi).max()
    ref = ref_oi @ G
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()
