"""GPU: the mixed-precision (BF16 hi+lo split) tetrapyd Gram against the FP64
path, the oracle and the exact Gauss-Legendre factorised FP64 Gram.  The
tolerances are the stated error of the path, max_mn |dG_mn| / sqrt(G_mm G_nn):
each operand is kept to ~2^-17 relative, and FP32 accumulation over one
segment (8192 points) adds the rest; measured 3.1e-7 at lmax=2000 with 1540
modes.  Small L, with few points per element, reaches a few 1e-6."""
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


@pytest.mark.parametrize("L,R,mod", [(60, 4, False), (131, 7, True), (300, 9, True)])
def test_mixed_vs_oracle_small(ctx, L, R, mod):
    ws = W.lspace(L, R, mod)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.MIXED)
    ref = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"])
    assert np.array_equal(G, G.T)
    e = rel_elem(G, ref)
    assert e < 1e-5, e   # few points per element here: less averaging of the split error


def test_mixed_lmax500_vs_fp64(ctx):
    c = W.WORKLOADS["lmax500_r10"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    G64 = ctx.gamma_tet(modalcmb.FP64)
    Gm = ctx.gamma_tet(modalcmb.MIXED)
    e = rel_elem(Gm, G64)
    assert e < 2e-6, e
    # the diagonal lo^2 correction removes the positive bias of the 3-product split
    d = np.diag(Gm) / np.diag(G64) - 1
    assert abs(d.mean()) < 2e-7


def test_mixed_lmax2000_vs_gl(ctx):
    """lmax=2000, 1540 modes (334,831,501 points): the full mixed Gram against
    the FP64 Gauss-Legendre factorisation of the same Gram."""
    c = W.WORKLOADS["lmax2000_r20"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    ref = ctx.gamma_gl()
    Gm = ctx.gamma_tet(modalcmb.MIXED)
    assert np.array_equal(Gm, Gm.T)
    e = rel_elem(Gm, ref)
    frob = float(np.linalg.norm(Gm - ref) / np.linalg.norm(ref))
    print(f"lmax 2000 mixed vs GL: max elem {e:.3e} frob {frob:.3e}")
    assert e < 1e-6, e
    assert frob < 1e-6, frob


def test_mixed_batched_dispatches(ctx):
    """A grid is split into several dispatches when it would exceed 2^32
    work-items (lmax=2500, 4960 modes needs it); forced here at a small size
    by capping the workgroups per dispatch, with short segments so that many
    dispatches run.  The result must match the single-dispatch Gram to FP64
    rounding of the (reordered) segment sums."""
    c = W.WORKLOADS["lmax500_r10"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    os.environ["MCMB_MIXED_SEGCH"] = "16"
    try:
        G1 = ctx.gamma_tet(modalcmb.MIXED)
        assert ctx.last_kernel_ms()[1] == 1      # one tile -> one dispatch
        os.environ["MCMB_MIXED_MAX_WG"] = "97"
        G2 = ctx.gamma_tet(modalcmb.MIXED)
        launches = ctx.last_kernel_ms()[1]
    finally:
        os.environ.pop("MCMB_MIXED_SEGCH", None)
        os.environ.pop("MCMB_MIXED_MAX_WG", None)
    assert launches > 10, launches
    assert rel_elem(G2, G1) < 1e-13
    G64 = ctx.gamma_tet(modalcmb.FP64)
    assert rel_elem(G2, G64) < 1e-6
