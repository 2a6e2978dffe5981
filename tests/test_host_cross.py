"""Host-side (no GPU) pieces of the cross-Gamma through the C ABI: the
per-sample weights of the x rules against the oracle's spline integrals,
and the 3D walk decomposition against the oracle, bit for bit."""
import ctypes as C

import numpy as np
import pytest

import modalcmb
import pyoracle as O


@pytest.mark.parametrize("n,seed", [(4, 0), (5, 1), (23, 2), (200, 3), (1000, 4)])
def test_spline_weights_match_oracle_integrals(n, seed):
    rng = np.random.default_rng(seed)
    x = np.sort(rng.uniform(0.0, 14.0, n))
    x = np.unique(x)
    for kind, name in ((modalcmb.SPLINE_FMM, "fmm"), (modalcmb.SPLINE_NATURAL, "natural")):
        c = modalcmb.spline_weights(kind, x)
        for _ in range(3):
            y = rng.standard_normal(len(x)) * np.exp(-x / 5)
            ref = O.spline_integral(x, y, name)
            assert float(c @ y) == pytest.approx(ref, rel=1e-11, abs=1e-13 * np.abs(y).sum())
    c = modalcmb.spline_weights(modalcmb.SPLINE_TRAPEZIUM, x)
    y = rng.standard_normal(len(x))
    assert float(c @ y) == pytest.approx(float(np.sum(np.diff(x) * (y[:-1] + y[1:]) / 2)), rel=1e-13)


def test_spline_weights_reject_short_or_unsorted_grids():
    with pytest.raises(modalcmb.ModalError):
        modalcmb.spline_weights(modalcmb.SPLINE_FMM, np.array([0.0, 1.0, 2.0]))
    with pytest.raises(modalcmb.ModalError):
        modalcmb.spline_weights(modalcmb.SPLINE_NATURAL, np.array([0.0, 2.0, 1.0, 3.0]))
    c = modalcmb.spline_weights(modalcmb.SPLINE_NATURAL, np.array([0.0, 2.0]))
    assert np.allclose(c, [1.0, 1.0])


class ThreadData(C.Structure):
    _fields_ = [(f, C.c_int) for f in ("i0", "j0", "k0", "i1", "j1", "k1", "xsize")] + \
               [(f, C.c_void_p) for f in ("mvec", "xvec", "yvec", "xdiff", "ixdiff", "intgrlvec", "plijkz", "bar")]


@pytest.mark.parametrize("lsize,rank,numranks,ws,nthreads",
                         [(21, 0, 1, 0.0, 1), (40, 0, 1, 0.0, 6), (40, 2, 3, 0.0, 4), (33, 1, 4, 0.25, 3),
                          (64, 6, 7, 0.0, 5)])
def test_decompose_gamma_3d_bit_exact(lsize, rank, numranks, ws, nthreads):
    lib = modalcmb.load_library()
    td = (ThreadData * nthreads)()
    lib.mcmb_decompose_gamma_TTT_3D(lsize, rank, numranks, C.c_double(ws), nthreads, td)
    got = np.array([[t.i0, t.j0, t.k0, t.i1, t.j1, t.k1] for t in td])
    assert np.array_equal(got, O.decompose_gamma_TTT_3D(lsize, rank, numranks, ws, nthreads))
