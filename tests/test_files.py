"""The reference's file names, the q-tilde files and the order-9 late basis
through the C ABI (host only, no GPU)."""
import ctypes as C
import os

import numpy as np
import pytest

import modalcmb

K = dict(eigen=0, ortho=1, **{"lambda": 2}, orthol=3, lambdal=4, gamma=5, proj=6, proj_size=7)


def name(kind, pol=0, directory="/data/", base="x", T=(2, 2500), E=(2, 2000), op=6, ol=9, amax=300):
    lib = modalcmb.load_library()
    buf = C.create_string_buffer(512)
    st = lib.mcmb_file_name(buf, 512, K[kind], pol, directory.encode(), base.encode(), T[0], T[1], E[0], E[1],
                            op, ol, amax)
    return st, buf.value.decode()


def test_file_names_follow_the_ini_pattern():
    # parameter.c:317-760 with suffix1..7 = T_lmax, E_lmax, alpha_max,
    # order_prim, order_late, T_lmin, E_lmin (244-257)
    assert name("gamma", 0, base="gamma") == (0, "/data/gamma_TTT_2-2500_6_9_300")
    assert name("gamma", 1, base="gamma") == (0, "/data/gamma_TTE_2-2500_2-2000_6_9_300")
    assert name("gamma", 2, base="gamma") == (0, "/data/gamma_TEE_2-2500_2-2000_6_9_300")
    assert name("gamma", 3, base="gamma") == (0, "/data/gamma_EEE_2-2000_6_9_300")
    assert name("orthol", 0, base="orthol") == (0, "/data/orthol_TTT_2-2500_9_300")
    assert name("lambdal", 2, base="lambdal") == (0, "/data/lambdal_TEE_2-2500_2-2000_9_300")
    assert name("eigen", base="eigen") == (0, "/data/eigen_2500_6_300")
    assert name("ortho", base="ortho") == (0, "/data/ortho_2500_6_300")
    assert name("lambda", base="lambda") == (0, "/data/lambda_2500_6_300")
    assert name("proj", 0, base="proj") == (0, "/data/proj_T_2500_6")
    assert name("proj", 1, base="proj") == (0, "/data/proj_E_2000_6")
    assert name("proj_size", 0, base="proj") == (0, "/data/proj_T_2500_size")
    # absolute base names
    assert name("gamma", 2, base="/abs/gam") == (0, "/abs/gam_TEE")
    assert name("proj_size", 1, base="/abs/p") == (0, "/abs/p_E_size")
    assert name("ortho", base="/abs/o") == (0, "/abs/o")
    assert name("orthol", 0, base="/abs/ol") == (0, "/abs/ol")
    assert name("orthol", 1, base="/abs/ol")[0] == -1     # the reference names only the TTT file
    # too small a buffer is an error, not a truncation
    lib = modalcmb.load_library()
    small = C.create_string_buffer(8)
    assert lib.mcmb_file_name(small, 8, 5, 0, b"/data/", b"gamma", 2, 2500, 2, 2000, 6, 9, 300) == -1


@pytest.mark.parametrize("order4", [0, 1])
def test_qtilde_files_round_trip(tmp_path, order4):
    lib = modalcmb.load_library()
    rng = np.random.default_rng(order4)
    npb, L, X = 5, 41, 17
    qt = rng.standard_normal((npb, L, X))
    lvec = np.arange(L, dtype=np.int32)
    xvec = np.linspace(0.0, 3.0, X)
    sf, df = str(tmp_path / "proj_T_40_size"), str(tmp_path / "proj_T_40_6")
    assert lib.mcmb_write_qtilde(sf.encode(), df.encode(), npb, order4, L, X, lvec.ctypes.data_as(C.c_void_p),
                                 xvec.ctypes.data_as(C.c_void_p), qt.ctypes.data_as(C.c_void_p)) == 0
    names = sorted(os.listdir(tmp_path))
    if order4:
        assert "proj_T_40_6_l1" in names and "proj_T_40_6_l2" in names and "proj_T_40_6_3" not in names
    else:
        assert "proj_T_40_6_4" in names
    # the layout: row 0 = (p, x grid), column 0 = multipoles
    a = np.fromfile(df + "_1", dtype=np.float64).reshape(L + 1, X + 1)
    assert a[0, 0] == 1.0 and np.array_equal(a[0, 1:], xvec) and np.array_equal(a[1:, 0], lvec)
    assert np.array_equal(a[1:, 1:], qt[1])
    assert np.array_equal(np.fromfile(sf, dtype=np.int32), [L + 1, X + 1])
    ls, xs = C.c_int(), C.c_int()
    assert lib.mcmb_read_qtilde(sf.encode(), df.encode(), npb, order4, 0, 0, C.byref(ls), C.byref(xs), None, None,
                                None) == 0
    assert (ls.value, xs.value) == (L, X)
    q2, l2, x2 = np.empty_like(qt), np.empty(L, dtype=np.int32), np.empty(X)
    assert lib.mcmb_read_qtilde(sf.encode(), df.encode(), npb, order4, L, X, C.byref(ls), C.byref(xs),
                                l2.ctypes.data_as(C.c_void_p), x2.ctypes.data_as(C.c_void_p),
                                q2.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(q2, qt) and np.array_equal(l2, lvec) and np.array_equal(x2, xvec)
    # capacities too small, and a missing file, are errors
    assert lib.mcmb_read_qtilde(sf.encode(), df.encode(), npb, order4, L - 1, X, C.byref(ls), C.byref(xs),
                                l2.ctypes.data_as(C.c_void_p), x2.ctypes.data_as(C.c_void_p),
                                q2.ctypes.data_as(C.c_void_p)) == -1
    assert lib.mcmb_read_qtilde(sf.encode(), df.encode(), npb + 1, 0, L, X, C.byref(ls), C.byref(xs),
                                l2.ctypes.data_as(C.c_void_p), x2.ctypes.data_as(C.c_void_p),
                                q2.ctypes.data_as(C.c_void_p)) == -5


def test_order9_late_basis():
    lib = modalcmb.load_library()
    Tmin, Tmax, p = 2, 300, 10
    tv = np.arange(Tmin, Tmax + 1, dtype=np.float64)
    out = np.empty((len(tv), p + 1))
    assert lib.mcmb_basis_late_order9(len(tv), p, C.c_double(Tmin), C.c_double(Tmax), tv.ctypes.data_as(C.c_void_p),
                                      out.ctypes.data_as(C.c_void_p)) == 0
    t = (tv - Tmin) / (Tmax - Tmin)
    for j in range(p - 1):   # shifted Legendre P_j(2t - 1)
        ref = np.polynomial.legendre.legval(2 * t - 1, np.eye(p - 1)[j])
        assert np.abs(out[:, j] - ref).max() < 1e-13
    x = tv * (tv + 1) / (Tmax * (Tmax + 1.0))
    w = ((2 * tv + 1) / (2 * Tmax + 1.0)) ** (1 / 6)
    assert np.allclose(out[:, p - 1], w / np.sqrt(x), rtol=1e-14, atol=0)
    assert np.allclose(out[:, p], w * np.sqrt(x), rtol=1e-14, atol=0)
