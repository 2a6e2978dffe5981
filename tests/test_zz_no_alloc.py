"""mcmb_gamma_gl_noalloc (the GL Gram without device allocation, scratch in a
per-GPU static array) against mcmb_gamma_gl and the oracle, through the C ABI:
two calls in a row on one context, the second on a smaller setup.  The
allocation-free entry always takes the direct node sum; mcmb_gamma_gl takes
the compressed pair basis where the pair products have one (both settings of
these tests), so the two agree bit for bit under MCMB_GL_PATH=direct and to
rounding otherwise."""
import ctypes as C
import os

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    return float((np.abs(A - B) / d).max())


def test_gamma_gl_noalloc_matches_allocating_entry_point():
    lib = modalcmb.load_library()
    with modalcmb.Context(0) as ctx:
        cases = [(300, 8, True), (150, 5, False)]       # second call: smaller input
        n_max = W.order_lex(8).shape[0]
        cap = n_max * n_max + 4096
        dev = ctx.device_alloc(8 * cap)
        try:
            prev = None
            for L, R, mod in cases:
                ws = W.lspace(L, R, mod)
                ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
                n = ws["order"].shape[0]
                st = lib.mcmb_gamma_gl_noalloc(ctx._h, 0, C.c_void_p(dev))
                assert st == 0, lib.mcmb_last_error().decode()
                buf = ctx.device_to_host(dev, np.empty(cap))
                G = buf[:n * n].reshape(n, n)
                os.environ["MCMB_GL_PATH"] = "direct"
                try:
                    ref = ctx.gamma_gl()
                finally:
                    del os.environ["MCMB_GL_PATH"]
                assert np.array_equal(G, ref)           # same kernels, same nodes
                assert rel_elem(G, ctx.gamma_gl()) < 1e-13
                assert rel_elem(G, O.gamma_gl(2, L, ws["q"], ws["v"], ws["order"])) < 1e-12
                if prev is not None:
                    # nothing written past the smaller call's n x n output
                    assert np.array_equal(buf[n * n:], prev[n * n:])
                prev = buf
            # a node table beyond the scratch's limit is refused, nothing run
            assert lib.mcmb_gamma_gl_noalloc(ctx._h, 5000, C.c_void_p(dev)) == -1
            assert b"limit" in lib.mcmb_last_error()
            assert lib.mcmb_gamma_gl_noalloc(ctx._h, 0, None) == -1
        finally:
            ctx.device_free(dev)
