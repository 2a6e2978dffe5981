"""CPU: file layouts of the hot-path outputs and order tables, and the
reference-named host helpers that need no device."""
import ctypes as C
import os

import numpy as np
import pytest

import modalcmb
import pyoracle as O


class ShapeParams(C.Structure):
    _fields_ = [("a1", C.c_double), ("a2", C.c_double), ("a3", C.c_double)]


def test_raw_roundtrip_and_layout(tmp_path):
    a = np.random.default_rng(1).standard_normal((7, 5))
    p = str(tmp_path / "gamma_TTT")
    modalcmb.write_raw(p, a)
    assert os.path.getsize(p) == a.size * 8                 # no header
    assert np.array_equal(np.fromfile(p, dtype=np.float64).reshape(7, 5), a)
    assert np.array_equal(modalcmb.read_raw(p, (7, 5)), a)


def test_raw_short_read_and_missing_file_are_errors(tmp_path):
    p = str(tmp_path / "short")
    modalcmb.write_raw(p, np.zeros(3))
    with pytest.raises(modalcmb.ModalError):
        modalcmb.read_raw(p, (4,))
    with pytest.raises(modalcmb.ModalError):
        modalcmb.read_raw(str(tmp_path / "none"), (1,))


def test_order_file_as_load_three_int(tmp_path):
    p = tmp_path / "order_dist.txt"
    p.write_text("# r s t\n0 0 0\n0 0 1\n\n1.0 2.9 3\n# trailing comment\n2\t2\t2\n")
    o = modalcmb.load_order(str(p))
    assert o.tolist() == [[0, 0, 0], [0, 0, 1], [1, 2, 3], [2, 2, 2]]   # fields truncated as (int)strtod
    lib = modalcmb.load_library()
    vals = np.zeros(3 * 16, dtype=np.int32)
    size = C.c_int(0)
    assert lib.load_three_int(str(p).encode(), vals.ctypes.data_as(C.c_void_p), C.byref(size)) == 1
    assert size.value == 4 and vals[:12].tolist() == sum(o.tolist(), [])
    assert lib.load_three_int(str(tmp_path / "none").encode(), vals.ctypes.data_as(C.c_void_p), C.byref(size)) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("0 0 0\nabc\n")
    with pytest.raises(modalcmb.ModalError):
        modalcmb.load_order(str(bad))


def test_shape3_matches_oracle_feature():
    lib = modalcmb.load_library()
    lib.shape3.restype = C.c_double
    lib.shape3.argtypes = [C.c_double, C.c_double, C.c_double, ShapeParams]
    lib.shape.restype = C.c_double
    lib.shape.argtypes = [C.c_double, C.c_double, C.c_double, ShapeParams]
    rng = np.random.default_rng(2)
    for _ in range(50):
        k1, k2, k3 = rng.uniform(1e-3, 0.2, 3)
        a1, a2 = rng.uniform(10, 500), rng.uniform(0, 100)
        sp = ShapeParams(a1, a2, 0.0)
        assert lib.shape3(k1, k2, k3, sp) == O.lib().or_shape3_feature(k1, k2, k3, a1, a2)
        assert lib.shape(k1, k2, k3, sp) == O.lib().or_feature(k1, k2, k3, a1, a2)
    assert lib.shape3(0.0, 0.1, 0.1, ShapeParams(1.0, 2.0, 0.0)) == 0.0


def test_helpers_without_bound_context_are_zero():
    lib = modalcmb.load_library()
    lib.mcmb_legacy_bind(None)
    assert lib.pijk(0, 1, 2, 3) == 0.0 and lib.plijk_TTT(0, 2, 3, 4) == 0.0
    assert lib.get_terms_prim() == 0 and lib.get_kmax() == 0.0
