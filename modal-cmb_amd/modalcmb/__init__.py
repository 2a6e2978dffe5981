"""Python binding of libmodalcmb (the MI355X Gamma / shape-projection library).

A thin ctypes layer over the C ABI declared in ``include/modal_cmb.h``.  All
compute runs in the HIP library; this module only moves host arrays in and out
and turns status codes into exceptions.  Importing it does not touch the GPU;
creating a :class:`Context` does, and raises if the library or a device is
missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libmodalcmb.so")

FP64 = 0
MIXED = 1
FP32 = 2
H2_EXACT = 0
H2_ASYMPTOTIC = 1
POL_TTT, POL_TTE, POL_TEE, POL_EEE = 0, 1, 2, 3

_STATUS = {
    0: "ok", -1: "bad argument", -2: "HIP runtime error", -3: "matrix not positive definite",
    -4: "bad mode order table", -5: "I/O error", -6: "out of memory", -7: "setup call missing",
}


class ModalError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}: {_STATUS.get(status, '?')}] {message}")
        self.status = status


class NotSPDError(ModalError):
    pass


_lib = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libmodalcmb.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise OSError(f"libmodalcmb.so not built at {path}: run make -C modal-cmb_amd")
        lib = C.CDLL(path)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        lib.mcmb_last_error.restype = C.c_char_p
        lib.mcmb_status_string.restype = C.c_char_p
        lib.mcmb_create.argtypes = [C.POINTER(vp), ip]
        lib.mcmb_destroy.argtypes = [vp]
        lib.mcmb_sync.argtypes = [vp]
        lib.mcmb_lspace_set.argtypes = [vp, ip, ip, ip, vp, vp, ip, vp, ip]
        lib.mcmb_gamma_tet.argtypes = [vp, ip, vp, vp]
        lib.mcmb_gamma_gl.argtypes = [vp, ip, vp, vp]
        lib.mcmb_gamma_gl_noalloc.argtypes = [vp, ip, vp]
        lib.mcmb_lspace_set_E.argtypes = [vp, ip, ip, ip, vp, vp, ip, vp, vp, vp]
        lib.mcmb_lspace_clear_E.argtypes = [vp]
        lib.mcmb_lspace_E_info.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.POINTER(C.c_long)]
        lib.mcmb_gamma_gl_pol.argtypes = [vp, ip, ip, vp, vp]
        lib.mcmb_qtilde_set_E.argtypes = [vp] + [vp] * 7 + [dp]
        lib.mcmb_qtilde_E_info.argtypes = [vp, C.POINTER(dp), C.POINTER(C.c_long)]
        lib.mcmb_gamma_cross_gl_pol.argtypes = [vp, ip, ip, ip, vp, vp]
        lib.plijk_TTE.restype = dp
        lib.plijk_TTE.argtypes = [ip, ip, ip, ip]
        lib.plijk_TEE.restype = dp
        lib.plijk_TEE.argtypes = [ip, ip, ip, ip]
        lib.plijk_EEE.restype = dp
        lib.plijk_EEE.argtypes = [ip, ip, ip, ip]
        for f in ("get_ijk_next_XXY", "check_ijk_min_XXY", "check_ijk_max_XXY"):
            n_int = 4 if f == "get_ijk_next_XXY" else 2
            getattr(lib, f).argtypes = [ip] * n_int + [C.POINTER(ip)] * 3
        lib.decompose_tetrapyd_XXY.argtypes = [vp, ip, ip, ip, ip, ip, ip, ip]
        lib.mcmb_gamma_tet_slice.argtypes = [vp, ip, ip, ip, vp, vp, C.POINTER(dp)]
        lib.mcmb_lspace_npoints.argtypes = [vp]
        lib.mcmb_lspace_npoints.restype = C.c_long
        lib.mcmb_lspace_points.argtypes = [vp, vp, C.c_long]
        lib.mcmb_lspace_points.restype = C.c_long
        lib.mcmb_lspace_points_at.argtypes = [vp, C.c_long, vp, C.c_long]
        lib.mcmb_lspace_points_at.restype = C.c_long
        lib.mcmb_gamma_tet_lrange.argtypes = [vp, ip, ip, ip, vp, vp]
        lib.mcmb_lspace_weights.argtypes = [vp, vp, C.c_long]
        lib.mcmb_lspace_weights.restype = C.c_long
        lib.mcmb_last_kernel_ms.argtypes = [vp, C.POINTER(dp), C.POINTER(ip)]
        lib.mcmb_last_phases.argtypes = [vp, ip, C.POINTER(dp), C.POINTER(C.c_char_p)]
        lib.mcmb_last_setup.argtypes = [vp, ip, C.POINTER(dp), C.POINTER(C.c_char_p)]
        lib.mcmb_last_basis_rank.argtypes = [vp, C.POINTER(ip)]
        lib.mcmb_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
        lib.mcmb_device_free.argtypes = [vp, vp]
        lib.mcmb_device_to_host.argtypes = [vp, vp, vp, C.c_size_t]
        lib.mcmb_device_copy.argtypes = [vp, vp, vp, C.c_size_t]
        lib.mcmb_orthonormalise.argtypes = [vp, ip, vp, vp, vp, vp, C.POINTER(ip)]
        lib.mcmb_cholesky_solve.argtypes = [vp, ip, vp, ip, vp, vp]
        lib.mcmb_eigen_sym.argtypes = [vp, ip, vp, vp, vp, C.POINTER(ip)]
        lib.mcmb_kuniform_set.argtypes = [vp, ip, ip, dp, ip, vp, ip, vp]
        lib.mcmb_gamma_post.argtypes = [vp, ip, vp, vp, vp, vp]
        lib.mcmb_write_raw.argtypes = [C.c_char_p, C.c_long, vp]
        lib.mcmb_read_raw.argtypes = [C.c_char_p, C.c_long, vp]
        lib.mcmb_load_order.argtypes = [C.c_char_p, vp, ip, C.POINTER(ip)]
        for f in ("pijk", "plijk_TTT"):
            getattr(lib, f).restype = dp
            getattr(lib, f).argtypes = [ip, ip, ip, ip]
        lib.get_kmax.restype = dp
        lib.mcmb_kuniform_info.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.POINTER(dp),
                                           C.POINTER(C.c_long)]
        lib.mcmb_kuniform_npoints.argtypes = [vp]
        lib.mcmb_kuniform_npoints.restype = C.c_long
        lib.mcmb_kuniform_weights.argtypes = [vp, vp, C.c_long]
        lib.mcmb_kuniform_weights.restype = C.c_long
        lib.mcmb_kuniform_gram.argtypes = [vp, ip, C.c_long, C.c_long, vp, vp]
        lib.mcmb_kuniform_project.argtypes = [vp, ip, vp, C.c_long, C.c_long, vp]
        lib.mcmb_orthonormalise_late.argtypes = [vp, ip, vp, ip, vp, vp, vp, vp, C.POINTER(ip)]
        for name in ("get_ijk_next_XXX", "get_ijk_next_prim"):
            getattr(lib, name).argtypes = [ip, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip)]
        lib.mcmb_gauss_legendre.argtypes = [ip, vp, vp]
        lib.mcmb_kspace_set.argtypes = [vp, ip, vp, vp, ip, vp, ip, vp]
        lib.mcmb_kspace_npoints.argtypes = [vp]
        lib.mcmb_kspace_npoints.restype = C.c_long
        lib.mcmb_kgram.argtypes = [vp, ip, vp, vp]
        lib.mcmb_project_fwd.argtypes = [vp, ip, ip, vp, vp, vp]
        lib.mcmb_project_bwd.argtypes = [vp, ip, ip, vp, vp, vp]
        lib.mcmb_qtilde_set.argtypes = [vp, ip, ip, ip, vp, vp, ip, vp, vp, vp, vp, dp]
        lib.mcmb_qtilde_info.argtypes = [vp] + [C.POINTER(ip)] * 4 + [C.POINTER(C.c_long)]
        lib.mcmb_gamma_cross_3d.argtypes = [vp, ip, ip, vp, vp]
        lib.mcmb_gamma_cross_gl.argtypes = [vp, ip, ip, vp, vp]
        lib.mcmb_spline_weights.argtypes = [ip, ip, vp, vp]
        lib.init_orthol_glint.restype = dp
        lib.mcmb_file_name.argtypes = [C.c_char_p, C.c_size_t, ip, ip, C.c_char_p, C.c_char_p] + [ip] * 7
        lib.mcmb_write_qtilde.argtypes = [C.c_char_p, C.c_char_p, ip, ip, ip, ip, vp, vp, vp]
        lib.mcmb_read_qtilde.argtypes = [C.c_char_p, C.c_char_p, ip, ip, ip, ip, C.POINTER(ip), C.POINTER(ip), vp,
                                         vp, vp]
        lib.mcmb_basis_late_order9.argtypes = [ip, ip, dp, dp, vp, vp]
        lib.permsix.restype = dp
        lib.permsix.argtypes = [ip, ip, ip]
        lib.calculate_geometric.restype = dp
        lib.calculate_geometric.argtypes = [ip, ip, ip]
        _lib = lib
    return _lib


EXPORTED_SYMBOLS = (
    "mcmb_last_error", "mcmb_status_string", "mcmb_create", "mcmb_destroy", "mcmb_sync",
    "mcmb_lspace_set", "mcmb_gamma_tet", "mcmb_gamma_gl", "mcmb_gamma_gl_noalloc", "mcmb_lspace_npoints",
    "mcmb_lspace_points", "mcmb_lspace_weights", "mcmb_orthonormalise", "mcmb_cholesky_solve",
    "mcmb_last_kernel_ms", "mcmb_last_phases", "mcmb_last_setup", "mcmb_last_basis_rank", "mcmb_device_alloc", "mcmb_device_free", "mcmb_device_to_host", "mcmb_device_copy", "mcmb_legacy_bind", "mcmb_legacy_status", "mcmb_lspace_info",
    "mcmb_gauss_legendre", "mcmb_gamma_tet_slice", "mcmb_gamma_tet_lrange", "mcmb_lspace_points_at", "mcmb_kspace_set", "mcmb_kspace_npoints",
    "mcmb_kgram", "mcmb_project_fwd", "mcmb_project_bwd", "mcmb_eigen_sym", "mcmb_orthonormalise_late",
    "mcmb_kuniform_set", "mcmb_kuniform_info", "mcmb_kuniform_npoints", "mcmb_kuniform_weights",
    "mcmb_kuniform_gram", "mcmb_kuniform_project", "calculate_ortho", "calculate_eigen",
    "mcmb_kuniform_tables", "mcmb_lspace_tables", "mcmb_gamma_post", "mcmb_write_raw", "mcmb_read_raw",
    "mcmb_load_order", "pijk", "plijk_TTT", "find_perm_prim", "get_terms_prim", "get_kmax", "shape", "shape3",
    "create_basis_prim", "load_three_int",
    "get_ijk_next_XXX", "check_ijk_min_XXX", "check_ijk_max_XXX", "decompose_tetrapyd_XXX",
    "get_ijk_next_prim", "check_ijk_min_prim", "check_ijk_max_prim", "decompose_tetrapyd_prim",
    "divide_tasks", "get_next_ijk_bound", "permsix", "calculate_geometric",
    "basis_functions_legendre", "gamma_3d_host", "calculate_orthol_3D", "calculate_orthol",
    "mcmb_qtilde_set", "mcmb_qtilde_info", "mcmb_qtilde_xvec", "mcmb_gamma_cross_3d", "mcmb_gamma_cross_gl",
    "mcmb_spline_weights", "mcmb_decompose_gamma_TTT_3D", "decompose_gamma_TTT_3D_mpi", "calculate_gamma_2d",
    "calculate_gamma", "init_gamma_glint", "precompute_gammaMij_T", "calculate_gammaMij_T", "sync_gammaMij",
    "get_qtilde_xsize", "get_qtilde_xvec", "get_qtilde_T_lsize", "init_orthol_glint", "calculate_orthoMij_T",
    "sync_orthoMij", "mcmb_lspace_set_E", "mcmb_lspace_clear_E", "mcmb_lspace_E_info", "mcmb_lspace_E_tables", "mcmb_gamma_gl_pol",
    "calculate_orthoMij_E", "find_perm_late_TTE", "find_perm_late_TEE", "find_perm_late_EEE", "plijk_TTE",
    "plijk_TEE", "plijk_EEE", "get_ijk_next_XXY", "check_ijk_min_XXY", "check_ijk_max_XXY",
    "decompose_tetrapyd_XXY", "mcmb_qtilde_set_E", "mcmb_qtilde_E_info", "mcmb_gamma_cross_gl_pol",
    "precompute_gammaMij_E", "calculate_gammaMij_E", "mcmb_file_name", "mcmb_write_qtilde", "mcmb_read_qtilde", "mcmb_basis_late_order9",
    "find_perm_late_TTT",
)


SPLINE_FMM, SPLINE_NATURAL, SPLINE_TRAPEZIUM = 0, 1, 2


def spline_weights(kind: int, x: np.ndarray) -> np.ndarray:
    """Per-sample weights of an x rule (host only)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    c = np.empty_like(x)
    _check(load_library().mcmb_spline_weights(kind, x.shape[0], _ptr(x), _ptr(c)))
    return c


def write_raw(path: str, a: np.ndarray) -> None:
    """array_write layout: raw native FP64, no header."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    _check(load_library().mcmb_write_raw(path.encode(), a.size, _ptr(a)))


def read_raw(path: str, shape) -> np.ndarray:
    a = np.empty(shape, dtype=np.float64)
    _check(load_library().mcmb_read_raw(path.encode(), a.size, _ptr(a)))
    return a


def load_order(path: str, cap: int = 1 << 20) -> np.ndarray:
    """Mode order table as load_three_int reads it."""
    buf = np.empty((cap, 3), dtype=np.int32)
    n = C.c_int(0)
    _check(load_library().mcmb_load_order(path.encode(), _ptr(buf), cap, C.byref(n)))
    return buf[:n.value].copy()


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _check(status: int) -> None:
    if status != 0:
        msg = _lib.mcmb_last_error().decode(errors="replace")
        if status == -3:
            raise NotSPDError(status, msg)
        raise ModalError(status, msg)


class Context:
    """One GPU, one stream; owns the device tables of the l-space setup."""

    def __init__(self, device: int = 0):
        lib = load_library()
        h = C.c_void_p()
        _check(lib.mcmb_create(C.byref(h), device))
        self._h = h
        self.n = 0

    def close(self) -> None:
        if self._h:
            _lib.mcmb_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- l-space Gram -------------------------------------------------------
    def lspace_set(self, lmin: int, lmax: int, q: np.ndarray, v: np.ndarray, order: np.ndarray,
                   h2_mode: int = H2_EXACT) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        order = np.ascontiguousarray(order, dtype=np.int32).reshape(-1, 3)
        if q.ndim != 2 or q.shape[0] != lmax + 1 or v.shape[0] != lmax + 1:
            raise ValueError("q must be [(lmax+1), ldq] and v [(lmax+1)]")
        self._keep = (q, v, order)
        _check(_lib.mcmb_lspace_set(self._h, lmin, lmax, q.shape[1], _ptr(q), _ptr(v),
                                    order.shape[0], _ptr(order), h2_mode))
        self.n = order.shape[0]

    def gamma_tet(self, precision: int = FP64) -> np.ndarray:
        g = np.empty((self.n, self.n), dtype=np.float64)
        _check(_lib.mcmb_gamma_tet(self._h, precision, _ptr(g), None))
        return g

    def gamma_tet_slice(self, precision: int, rank: int, nranks: int):
        """Partial Gram of one contiguous share of the tetrapyd; returns
        (gamma_part, work_fraction)."""
        g = np.empty((self.n, self.n), dtype=np.float64)
        frac = C.c_double()
        _check(_lib.mcmb_gamma_tet_slice(self._h, precision, rank, nranks, _ptr(g), None, C.byref(frac)))
        return g, frac.value

    def gamma_tet_lrange(self, precision: int, l1_lo: int, l1_hi: int) -> np.ndarray:
        """Partial Gram over the points with l1 in [l1_lo, l1_hi)."""
        g = np.empty((self.n, self.n), dtype=np.float64)
        _check(_lib.mcmb_gamma_tet_lrange(self._h, precision, l1_lo, l1_hi, _ptr(g), None))
        return g

    # ---- cross-Gamma against q-tilde --------------------------------------
    def qtilde_set(self, qt: np.ndarray, xvec: np.ndarray, order_prim: np.ndarray, cl: np.ndarray,
                   nl: np.ndarray, bl: np.ndarray, deltaphi: float) -> None:
        """qt[p][l][x] (read_qtilde layout, l = multipole 0..lmax_q)."""
        qt = np.ascontiguousarray(qt, dtype=np.float64)
        np_, nlq, nx = qt.shape
        xvec = np.ascontiguousarray(xvec, dtype=np.float64)
        op = np.ascontiguousarray(order_prim, dtype=np.int32).reshape(-1, 3)
        sp = [np.ascontiguousarray(a, dtype=np.float64) for a in (cl, nl, bl)]
        _check(_lib.mcmb_qtilde_set(self._h, np_, nlq - 1, nx, _ptr(qt), _ptr(xvec), op.shape[0], _ptr(op),
                                    _ptr(sp[0]), _ptr(sp[1]), _ptr(sp[2]), float(deltaphi)))
        self.nprim = op.shape[0]

    def qtilde_set_E(self, qt_E: np.ndarray, cl_TE, cl_EE, nl_TE, nl_EE, bl_TE, bl_EE, fsky_ratio: float) -> None:
        """E-mode q-tilde of the T q-tilde's shape and the TE / EE spectra per multipole."""
        qt_E = np.ascontiguousarray(qt_E, dtype=np.float64)
        sp = [np.ascontiguousarray(a, dtype=np.float64) for a in (cl_TE, cl_EE, nl_TE, nl_EE, bl_TE, bl_EE)]
        self._keepQE = (qt_E, sp)
        _check(_lib.mcmb_qtilde_set_E(self._h, _ptr(qt_E), *(_ptr(a) for a in sp), float(fsky_ratio)))

    def gamma_cross_gl_pol(self, pol: int, spline: int = SPLINE_FMM, npt: int = 0) -> np.ndarray:
        n = self.n if pol == POL_TTT else self.nE
        g = np.empty((n, self.nprim))
        _check(_lib.mcmb_gamma_cross_gl_pol(self._h, pol, spline, npt, _ptr(g), None))
        return g

    def gamma_cross_3d(self, rank: int = 0, nranks: int = 1) -> np.ndarray:
        g = np.empty((self.n, self.nprim))
        _check(_lib.mcmb_gamma_cross_3d(self._h, rank, nranks, _ptr(g), None))
        return g

    def gamma_cross_gl(self, spline: int = 0, npt: int = 0) -> np.ndarray:
        g = np.empty((self.n, self.nprim))
        _check(_lib.mcmb_gamma_cross_gl(self._h, spline, npt, _ptr(g), None))
        return g

    def gamma_gl(self, npt: int = 0) -> np.ndarray:
        g = np.empty((self.n, self.n), dtype=np.float64)
        _check(_lib.mcmb_gamma_gl(self._h, npt, _ptr(g), None))
        return g

    def lspace_set_E(self, E_lmin: int, E_lmax: int, qE: np.ndarray, vE: np.ndarray, order_TTE: np.ndarray,
                     order_TEE: np.ndarray, order_EEE: np.ndarray) -> None:
        """E field of the late basis, qE[l][ldq] / vE[l] for l = 0..E_lmax, and the
        TTE (T, T, E), TEE (T, E, E), EEE mode tables (one row count)."""
        qE = np.ascontiguousarray(qE, dtype=np.float64)
        vE = np.ascontiguousarray(vE, dtype=np.float64)
        tabs = [np.ascontiguousarray(t, dtype=np.int32).reshape(-1, 3) for t in (order_TTE, order_TEE, order_EEE)]
        if qE.ndim != 2 or qE.shape[0] != E_lmax + 1 or vE.shape[0] != E_lmax + 1:
            raise ValueError("qE must be [(E_lmax+1), ldq] and vE [(E_lmax+1)]")
        if len({t.shape[0] for t in tabs}) != 1:
            raise ValueError("the three mode tables must have one row count")
        self._keepE = (qE, vE, tabs)
        _check(_lib.mcmb_lspace_set_E(self._h, E_lmin, E_lmax, qE.shape[1], _ptr(qE), _ptr(vE), tabs[0].shape[0],
                                      _ptr(tabs[0]), _ptr(tabs[1]), _ptr(tabs[2])))
        self.nE = tabs[0].shape[0]

    def lspace_clear_E(self) -> None:
        _check(_lib.mcmb_lspace_clear_E(self._h))
        self.nE = 0

    def gamma_gl_pol(self, pol: int, npt: int = 0) -> np.ndarray:
        """Component pol (POL_TTT .. POL_EEE) of the polarised GL Gram."""
        n = self.n if pol == POL_TTT else self.nE
        g = np.empty((n, n), dtype=np.float64)
        _check(_lib.mcmb_gamma_gl_pol(self._h, pol, npt, _ptr(g), None))
        return g

    def gamma_tet_device(self, precision: int, dev_ptr: int) -> None:
        _check(_lib.mcmb_gamma_tet(self._h, precision, None, C.c_void_p(dev_ptr)))

    def device_alloc(self, nbytes: int) -> int:
        """Device buffer on the context's GPU (an integer address)."""
        p = C.c_void_p()
        _check(_lib.mcmb_device_alloc(self._h, nbytes, C.byref(p)))
        return int(p.value or 0)

    def device_free(self, dev_ptr: int) -> None:
        _check(_lib.mcmb_device_free(self._h, C.c_void_p(dev_ptr)))

    def device_to_host(self, dev_ptr: int, out: np.ndarray) -> np.ndarray:
        _check(_lib.mcmb_device_to_host(self._h, _ptr(out), C.c_void_p(dev_ptr), out.nbytes))
        return out

    def device_copy(self, dst: int, src: int, nbytes: int) -> None:
        """Plain copy kernel of nbytes (multiple of 16) on the device."""
        _check(_lib.mcmb_device_copy(self._h, C.c_void_p(dst), C.c_void_p(src), nbytes))

    def npoints(self) -> int:
        r = _lib.mcmb_lspace_npoints(self._h)
        if r < 0:
            _check(int(r))
        return int(r)

    def points(self, cap: int | None = None) -> np.ndarray:
        n = self.npoints()
        cap = n if cap is None else min(cap, n)
        out = np.empty((cap, 3), dtype=np.int32)
        r = _lib.mcmb_lspace_points(self._h, _ptr(out), cap)
        if r < 0:
            _check(int(r))
        return out

    def points_at(self, p0: int, cap: int) -> np.ndarray:
        """Points p0 .. p0 + cap - 1 of the device enumeration (fewer at its end)."""
        n = self.npoints()
        m = max(0, min(cap, n - p0))
        out = np.empty((m, 3), dtype=np.int32)
        r = _lib.mcmb_lspace_points_at(self._h, p0, _ptr(out), m)
        if r < 0:
            _check(int(r))
        return out

    def weights(self, cap: int | None = None) -> np.ndarray:
        n = self.npoints()
        cap = n if cap is None else min(cap, n)
        out = np.empty(cap, dtype=np.float64)
        r = _lib.mcmb_lspace_weights(self._h, _ptr(out), cap)
        if r < 0:
            _check(int(r))
        return out

    def last_phases(self) -> list[tuple[str, float]]:
        """(name, ms) per phase of the last call that records phases."""
        cap = 16
        ms = (C.c_double * cap)()
        names = (C.c_char_p * cap)()
        k = _lib.mcmb_last_phases(self._h, cap, ms, names)
        if k < 0:
            _check(k)
        return [(names[i].decode(), float(ms[i])) for i in range(min(k, cap))]

    def last_setup(self) -> list[tuple[str, float]]:
        """(name, host wall ms) per part of the last separable Gram call."""
        cap = 20
        ms = (C.c_double * cap)()
        names = (C.c_char_p * cap)()
        k = _lib.mcmb_last_setup(self._h, cap, ms, names)
        if k < 0:
            _check(k)
        return [(names[i].decode(), float(ms[i])) for i in range(min(k, cap))]

    def last_basis_rank(self) -> int:
        """Basis columns of the last separable Gram's contractions (0: the
        pair products themselves; -1: another contraction order)."""
        r = C.c_int()
        _check(_lib.mcmb_last_basis_rank(self._h, C.byref(r)))
        return r.value

    def last_kernel_ms(self) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int()
        _check(_lib.mcmb_last_kernel_ms(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def sync(self) -> None:
        _check(_lib.mcmb_sync(self._h))

    # ---- k-space Gram and projection ---------------------------------------
    def kspace_set(self, kn: np.ndarray, kw: np.ndarray, q: np.ndarray, order: np.ndarray) -> None:
        kn = np.ascontiguousarray(kn, dtype=np.float64)
        kw = np.ascontiguousarray(kw, dtype=np.float64)
        q = np.ascontiguousarray(q, dtype=np.float64)
        order = np.ascontiguousarray(order, dtype=np.int32).reshape(-1, 3)
        self._kkeep = (kn, kw, q, order)
        _check(_lib.mcmb_kspace_set(self._h, kn.shape[0], _ptr(kn), _ptr(kw), q.shape[1], _ptr(q),
                                    order.shape[0], _ptr(order)))
        self.nk = order.shape[0]

    def kspace_npoints(self) -> int:
        r = _lib.mcmb_kspace_npoints(self._h)
        if r < 0:
            _check(int(r))
        return int(r)

    def kgram(self, precision: int = FP64) -> np.ndarray:
        g = np.empty((self.nk, self.nk), dtype=np.float64)
        _check(_lib.mcmb_kgram(self._h, precision, _ptr(g), None))
        return g

    def project_fwd(self, theta: np.ndarray, precision: int = FP64, derivatives: bool = True):
        """alpha[b][n] and (if derivatives) dalpha[b][2][n] (d/domega, d/dphi)."""
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1, 2))
        nt = th.shape[0]
        a = np.empty((nt, self.nk))
        d = np.empty((nt, 2, self.nk)) if derivatives else None
        _check(_lib.mcmb_project_fwd(self._h, precision, nt, _ptr(th), _ptr(a), _ptr(d)))
        return (a, d) if derivatives else a

    def project_bwd(self, theta: np.ndarray, gbar: np.ndarray, precision: int = FP64) -> np.ndarray:
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1, 2))
        g = np.ascontiguousarray(np.asarray(gbar, dtype=np.float64).reshape(th.shape[0], self.nk))
        out = np.empty((th.shape[0], 2))
        _check(_lib.mcmb_project_bwd(self._h, precision, th.shape[0], _ptr(th), _ptr(g), _ptr(out)))
        return out

    # ---- orthonormalisation --------------------------------------------------
    def orthonormalise(self, gamma: np.ndarray):
        """Return (ortho, lambda, norm) as ortho_prim_mpi.c:198-284 leaves them."""
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        n = g.shape[0]
        ortho = np.empty_like(g)
        lam = np.empty_like(g)
        norm = np.empty(n)
        bad = C.c_int(-1)
        st = _lib.mcmb_orthonormalise(self._h, n, _ptr(g), _ptr(ortho), _ptr(lam), _ptr(norm),
                                      C.byref(bad))
        if st == -3:
            raise NotSPDError(st, f"{_lib.mcmb_last_error().decode()} (pivot {bad.value})")
        _check(st)
        return ortho, lam, norm

    def eigen_sym(self, a: np.ndarray, vectors: bool = True):
        """(evals descending, evecs[:, k] or None, sweeps) on the GPU."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        ev = np.empty(n)
        V = np.empty_like(a) if vectors else None
        sw = C.c_int(0)
        _check(_lib.mcmb_eigen_sym(self._h, n, _ptr(a), _ptr(ev), _ptr(V), C.byref(sw)))
        return ev, V, sw.value

    def orthonormalise_late(self, gamma: np.ndarray, pca_flag: int = 1):
        """(r, lambda, lambdainv, norm, evals) as ortho_late_mpi.c:752-842 leaves them."""
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        n = g.shape[0]
        lam, linv, norm, ev = np.empty_like(g), np.empty_like(g), np.empty(n), np.empty(n)
        r = C.c_int(0)
        st = _lib.mcmb_orthonormalise_late(self._h, n, _ptr(g), pca_flag, _ptr(lam), _ptr(linv), _ptr(norm),
                                           _ptr(ev), C.byref(r))
        if st == -3:
            raise NotSPDError(st, _lib.mcmb_last_error().decode())
        _check(st)
        return r.value, lam, linv, norm, ev

    # ---- uniform k-grid (calculate_ortho / calculate_eigen) -------------------
    def kuniform_set(self, mn: int, mx: int, kmax: float, q: np.ndarray, order: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        order = np.ascontiguousarray(order, dtype=np.int32).reshape(-1, 3)
        self._ukeep = (q, order)
        _check(_lib.mcmb_kuniform_set(self._h, mn, mx, kmax, q.shape[1], _ptr(q), order.shape[0], _ptr(order)))
        self.nu = order.shape[0]

    def kuniform_npoints(self) -> int:
        r = _lib.mcmb_kuniform_npoints(self._h)
        if r < 0:
            _check(int(r))
        return int(r)

    def kuniform_weights(self) -> np.ndarray:
        n = self.kuniform_npoints()
        w = np.empty(n)
        r = _lib.mcmb_kuniform_weights(self._h, _ptr(w), n)
        if r < 0:
            _check(int(r))
        return w

    def kuniform_gram(self, precision: int = FP64, p0: int = 0, p1: int = -1) -> np.ndarray:
        g = np.empty((self.nu, self.nu))
        _check(_lib.mcmb_kuniform_gram(self._h, precision, p0, p1, _ptr(g), None))
        return g

    def kuniform_project(self, theta: np.ndarray, p0: int = 0, p1: int = -1) -> np.ndarray:
        th = np.ascontiguousarray(np.asarray(theta, dtype=np.float64).reshape(-1, 2))
        a = np.empty((th.shape[0], self.nu))
        _check(_lib.mcmb_kuniform_project(self._h, th.shape[0], _ptr(th), p0, p1, _ptr(a)))
        return a

    def gamma_post(self, lam: np.ndarray, gamma: np.ndarray):
        """(orthoinv = lambda^T lambda, orthoinv @ gamma) as gamma_mpi.c:260-295, 801-840."""
        lam = np.ascontiguousarray(lam, dtype=np.float64)
        g = np.ascontiguousarray(gamma, dtype=np.float64)
        oi, out = np.empty_like(g), np.empty_like(g)
        _check(_lib.mcmb_gamma_post(self._h, g.shape[0], _ptr(lam), _ptr(g), _ptr(oi), _ptr(out)))
        return oi, out

    def cholesky_solve(self, ortho: np.ndarray, alpha: np.ndarray) -> np.ndarray:
        f = np.ascontiguousarray(ortho, dtype=np.float64)
        n = f.shape[0]
        a = np.asfortranarray(np.asarray(alpha, dtype=np.float64).reshape(n, -1))
        out = np.empty_like(a, order="F")
        _check(_lib.mcmb_cholesky_solve(self._h, n, _ptr(f), a.shape[1], _ptr(a), _ptr(out)))
        return out.reshape(np.shape(alpha))
