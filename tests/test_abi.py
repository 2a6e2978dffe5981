"""The C-ABI library loads and exports every symbol include/modal_cmb.h
declares; its host-side (no GPU) reference entry points match the oracle
bit-exactly.  CPU only: no compute call touches a GPU here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import modalcmb
import pyoracle as O
from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "modal_cmb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:int|long|void|double|char)\s*\*?\s*(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = modalcmb.load_library()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(modalcmb.EXPORTED_SYMBOLS) <= set(names)


def _walk_lib(lmin, lmax):
    lib = modalcmb.load_library()
    i, j, k = C.c_int(lmin), C.c_int(lmin), C.c_int(lmin)
    lib.check_ijk_min_XXX(lmin, C.byref(i), C.byref(j), C.byref(k))
    out = [(i.value, j.value, k.value)]
    while lib.get_ijk_next_XXX(lmax, C.byref(i), C.byref(j), C.byref(k)):
        out.append((i.value, j.value, k.value))
    out.append((i.value, j.value, k.value))
    return out


@pytest.mark.parametrize("lmin,lmax", [(2, 10), (2, 11), (0, 9), (3, 17), (2, 60), (7, 41)])
def test_host_walk_bit_exact(lmin, lmax):
    assert _walk_lib(lmin, lmax) == [tuple(p) for p in O.walk_XXX(lmin, lmax)]


def test_host_decompositions_bit_exact():
    lib = modalcmb.load_library()
    for (rank, nranks, nthreads, lmin, lmax) in [(0, 1, 4, 2, 50), (2, 3, 5, 2, 77), (1, 2, 3, 0, 20)]:
        out = np.zeros((nthreads, 4), dtype=np.int32)
        lib.decompose_tetrapyd_XXX(out.ctypes.data_as(C.c_void_p), rank, nranks, nthreads, lmin, lmax)
        assert np.array_equal(out, O.decompose_XXX(rank, nranks, nthreads, lmin, lmax))
        lib.decompose_tetrapyd_prim(out.ctypes.data_as(C.c_void_p), rank, nranks, nthreads, lmin, lmax)
        assert np.array_equal(out, O.decompose_prim(rank, nranks, nthreads, lmin, lmax))


def test_host_prim_walk_bit_exact():
    lib = modalcmb.load_library()
    mx = 14
    i, j, k = C.c_int(0), C.c_int(0), C.c_int(0)
    lib.check_ijk_min_prim(0, C.byref(i), C.byref(j), C.byref(k))
    out = [(0, 0, 0)]
    while lib.get_ijk_next_prim(mx, C.byref(i), C.byref(j), C.byref(k)):
        out.append((i.value, j.value, k.value))
    out.append((i.value, j.value, k.value))
    assert out == [tuple(p) for p in O.walk_prim(0, mx)]


def test_host_block_bounds_bit_exact():
    lib = modalcmb.load_library()
    rng = np.random.default_rng(11)
    for lsize in (31, 64):
        pts = O.walk_XXX(2, lsize - 1)
        for _ in range(40):
            a = int(rng.integers(0, len(pts) - 1))
            i0, j0, k0 = pts[a].tolist()
            mb = int(rng.integers(0, 200))
            r = [C.c_int() for _ in range(4)]
            lib.mcmb_next_ijk_bound(lsize, i0, j0, k0, C.byref(r[0]), C.byref(r[1]), C.byref(r[2]),
                                    lsize, lsize, lsize, C.byref(r[3]), mb)
            assert tuple(x.value for x in r) == O.next_ijk_bound(lsize, i0, j0, k0, lsize, lsize, lsize, mb)


def test_host_weights_and_basis_match_oracle():
    lib = modalcmb.load_library()
    for t in [(2, 2, 2), (3, 4, 5), (100, 200, 300), (2, 3, 4), (2, 3, 9), (700, 800, 1499)]:
        assert lib.calculate_geometric(*t) == O.lib().or_geometric_asym(*t)
        assert lib.permsix(*t) == O.lib().or_permsix(*t)
    vals = np.linspace(0, 2000, 2001)
    rows = (C.POINTER(C.c_double) * 2001)()
    tab = np.empty((2001, 30))
    for r in range(2001):
        rows[r] = tab[r].ctypes.data_as(C.POINTER(C.c_double))
    lib.basis_functions_legendre(rows, 2001, 29, C.c_double(0.0), C.c_double(2000.0),
                                 vals.ctypes.data_as(C.c_void_p))
    assert np.array_equal(tab, O.basis_legendre(vals, 29, 0.0, 2000.0))


def test_host_gauss_legendre_matches_oracle():
    from modalcmb import workloads as W
    for n in (7, 200, 3751):
        x, w = W.gauss_legendre(n)
        xo, wo = O.gauss_legendre(n)
        assert np.abs(x - xo).max() <= 2.3e-16 and np.abs(w / wo - 1).max() < 1e-14


def test_context_without_gpu_fails_loudly():
    import subprocess, sys
    code = ("import sys; sys.path.insert(0, %r); import modalcmb\n"
            "try:\n    modalcmb.Context(0)\nexcept modalcmb.ModalError as e:\n    print('ERR', e.status)\n"
            "else:\n    print('OK')\n") % os.path.join(ROOT, "modal-cmb_amd")
    env = dict(os.environ, HIP_VISIBLE_DEVICES="-1", ROCR_VISIBLE_DEVICES="-1")
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
    assert out.stdout.strip().startswith("ERR"), out.stdout + out.stderr
