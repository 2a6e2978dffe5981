"""Synthetic inputs of the benchmark and parity workloads.

All inputs are generated deterministically (no data files):

* ``lspace(L, R, modulated)`` -- multipoles lmin=2..L, 1D basis
  q_r(l) = P_r(2l/L - 1) for l = 0..L (shifted Legendre by the three-term
  recurrence of basis_functions.c:88-107 with min=0, max=L), per-leg weight
  v(l) = 1/C_l with C_l = 1000/(l(l+1)) (times (1 + 0.1 sin(l/30)) when
  modulated), and the lexicographic order table r<=s<=t<R.
* ``kspace(N, R)`` -- N-node Gauss-Legendre grid on [2e-4, 0.2] Mpc^-1 per
  side, basis P_r(2k/kmax - 1) (create_basis_prim passes min=0,
  fixed_data.c:1688), same order table.
* ``shape_batch(B, seed)`` -- B (omega, phi) pairs from splitmix64:
  omega ~ U[10, 500] Mpc, phi ~ U[0, 2 pi).
"""
from __future__ import annotations

import numpy as np

KMIN, KMAX = 2e-4, 0.2

# l-space Grams (lmax L, basis width R) and k-space projections (N GL nodes per
# side, basis width R), by name
WORKLOADS = {
    "lmax500_r10": dict(L=500, R=10, modulated=False),
    "lmax2000_r20": dict(L=2000, R=20, modulated=True),
    "lmax2500_r30": dict(L=2500, R=30, modulated=True),
    "gl200_r20": dict(N=200, R=20, omega=150.0, phi=0.7),
    "gl300_r30_batch256": dict(N=300, R=30, batch=256, seed=20251202),
}


def legendre_table(x01: np.ndarray, pmax: int) -> np.ndarray:
    """Rows P_0..P_pmax at 2*x01 - 1 by the recurrence of basis_functions.c:100-104
    (same operation order, so the table is bit-identical to the reference's)."""
    x01 = np.asarray(x01, dtype=np.float64)
    a = np.empty((x01.shape[0], pmax + 1), dtype=np.float64)
    a[:, 0] = 1.0
    if pmax > 0:
        a[:, 1] = 2.0 * x01 - 1.0
    for j in range(2, pmax + 1):
        a[:, j] = ((2.0 * j - 1.0) * a[:, 1] * a[:, j - 1] - (j - 1.0) * a[:, j - 2]) / (1.0 * j)
    return a


def order_lex(R: int) -> np.ndarray:
    """(r, s, t), r <= s <= t < R, lexicographic (the explicit order file)."""
    return np.array([(r, s, t) for r in range(R) for s in range(r, R) for t in range(s, R)],
                    dtype=np.int32)


def lspace(L: int, R: int, modulated: bool = False, lmin: int = 2):
    ls = np.arange(L + 1, dtype=np.float64)
    q = legendre_table((ls - 0.0) / (float(L) - 0.0), R - 1)
    v = np.zeros(L + 1)
    cl = 1000.0 / (ls[1:] * (ls[1:] + 1.0))
    if modulated:
        cl = cl * (1.0 + 0.1 * np.sin(ls[1:] / 30.0))
    v[1:] = 1.0 / cl
    return dict(lmin=lmin, lmax=L, q=q, v=v, order=order_lex(R))


def gauss_legendre(n: int):
    """n-node Gauss-Legendre rule from the library's host routine (numpy's
    leggauss loses ~1e-9 in the end weights at n ~ 10^3)."""
    import ctypes as C
    from . import load_library
    x, w = np.empty(n), np.empty(n)
    st = load_library().mcmb_gauss_legendre(n, x.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p))
    if st != 0:
        raise ValueError(f"mcmb_gauss_legendre({n}) failed: {st}")
    return x, w


def kspace(N: int, R: int):
    x, w = gauss_legendre(N)
    kn = 0.5 * (KMAX - KMIN) * x + 0.5 * (KMAX + KMIN)
    kw = 0.5 * (KMAX - KMIN) * w
    q = legendre_table((kn - 0.0) / (KMAX - 0.0), R - 1)
    return dict(N=N, kn=kn, kw=kw, q=q, order=order_lex(R))


def _splitmix64(state: int):
    mask = (1 << 64) - 1
    while True:
        state = (state + 0x9E3779B97F4A7C15) & mask
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        yield z ^ (z >> 31)


def shape_batch(B: int, seed: int = 20251202) -> np.ndarray:
    """(omega, phi) pairs: omega ~ U[10,500] Mpc, phi ~ U[0, 2 pi), splitmix64(seed)."""
    g = _splitmix64(seed)
    out = np.empty((B, 2))
    for b in range(B):
        u1 = (next(g) >> 11) * (1.0 / (1 << 53))
        u2 = (next(g) >> 11) * (1.0 / (1 << 53))
        out[b, 0] = 10.0 + 490.0 * u1
        out[b, 1] = 2.0 * np.pi * u2
    return out
