"""GPU: the mixed-precision (BF16 hi+lo split) tetrapyd Gram against the FP64
path and the oracle.  Tolerances are the split's stated error: each operand is
kept to ~2^-17 relative, so Gamma_mn agrees with FP64 to about 1e-6 of
sqrt(Gamma_mm Gamma_nn) at L=500 (measured ~3e-7 in emulation); small L,
with few points per element, reaches a few 1e-6."""
import numpy as np
import pytest

import modalcmb
import pyoracle as O
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    return float((np.abs(A - B) / d).max())


@pytest.mark.parametrize("L,R,mod", [(60, 4, False), (131, 7, True), (300, 9, True)])
def test_mixed_vs_oracle_small(ctx, L, R, mod):
    ws = W.lspace(L, R, mod)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.MIXED)
    ref = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"])
    assert np.array_equal(G, G.T)
    e = rel_elem(G, ref)
    assert e < 1e-5, e   # few points per element here: less averaging of the split error


def test_mixed_config1_vs_fp64(ctx):
    c = W.CONFIGS[1]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    G64 = ctx.gamma_tet(modalcmb.FP64)
    Gm = ctx.gamma_tet(modalcmb.MIXED)
    e = rel_elem(Gm, G64)
    assert e < 2e-6, e
    # the diagonal lo^2 correction removes the positive bias of the 3-product split
    d = np.diag(Gm) / np.diag(G64) - 1
    assert abs(d.mean()) < 2e-7
