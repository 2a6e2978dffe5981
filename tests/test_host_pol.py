"""The library's mixed XXY walk (host C, tetrapyd_tools.h:81-84) against the
oracle, bit for bit, including the configurations where the reference's
walk is not the mixed tetrapyd (odd limits, unequal ranges)."""
import ctypes as C

import numpy as np
import pytest

import modalcmb
import pyoracle as O

CONFIGS = [(2, 2, 20, 20), (2, 2, 9, 9), (2, 3, 20, 20), (2, 2, 30, 17), (2, 2, 17, 30), (2, 4, 25, 25),
           (3, 2, 21, 26), (5, 2, 12, 40), (2, 10, 20, 20), (2, 2, 64, 64)]


def lib_walk(lminX, lminY, lmaxX, lmaxY, n):
    lib = modalcmb.load_library()
    i, j, k = C.c_int(lminX), C.c_int(lminX), C.c_int(lminY)
    lib.check_ijk_min_XXY(lminX, lminY, C.byref(i), C.byref(j), C.byref(k))
    out = np.empty((n, 3), dtype=np.int32)
    for p in range(n):
        out[p] = (i.value, j.value, k.value)
        if p + 1 < n:
            more = lib.get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, C.byref(i), C.byref(j), C.byref(k))
            assert more == (p + 2 < n)
    return out


@pytest.mark.parametrize("lims", CONFIGS)
def test_xxy_walk_bit_exact(lims):
    ref = O.walk_XXY(*lims)
    assert np.array_equal(lib_walk(*lims, len(ref)), ref)


@pytest.mark.parametrize("lims", CONFIGS[:6])
@pytest.mark.parametrize("rank,nranks,nthreads", [(0, 1, 1), (0, 1, 6), (2, 3, 4), (3, 4, 5)])
def test_decompose_xxy_bit_exact(lims, rank, nranks, nthreads):
    lib = modalcmb.load_library()
    blk = np.zeros((nthreads, 4), dtype=np.int32)
    lib.decompose_tetrapyd_XXY(blk.ctypes.data_as(C.c_void_p), rank, nranks, nthreads, *lims)
    assert np.array_equal(blk, O.decompose_XXY(rank, nranks, nthreads, *lims))


def test_xxy_checks_bit_exact():
    lib = modalcmb.load_library()
    rng = np.random.default_rng(4)
    for _ in range(3000):
        lminX, lminY = (int(x) for x in rng.integers(0, 12, 2))
        lmaxX, lmaxY = lminX + int(rng.integers(1, 30)), lminY + int(rng.integers(1, 30))
        p = tuple(int(x) for x in rng.integers(0, 45, 3))
        for name, a, b, ref in (("check_ijk_min_XXY", lminX, lminY, O.check_min_XXY),
                                ("check_ijk_max_XXY", lmaxX, lmaxY, O.check_max_XXY)):
            i, j, k = (C.c_int(x) for x in p)
            getattr(lib, name)(a, b, C.byref(i), C.byref(j), C.byref(k))
            assert (i.value, j.value, k.value) == ref(a, b, p), (name, a, b, p)
        i, j, k = (C.c_int(x) for x in p)
        more = lib.get_ijk_next_XXY(lminX, lminY, lmaxX, lmaxY, C.byref(i), C.byref(j), C.byref(k))
        assert (more, (i.value, j.value, k.value)) == O.next_XXY(lminX, lminY, lmaxX, lmaxY, p)
