"""The compressed pair basis of the separable Gram, on the CPU.

The separable Gram contracts over an orthonormal basis P of the span of the
pair products U_xy(l) = q_x(l) q_y(l) when their numerical rank r is at most
half the NP = R(R+1)/2 pairs, then expands the symmetrised tensor back to
pairs.  These tests pin what that rests on, independently of the GPU code:

* the rank of the pair products of the workloads' Legendre bases is 2R - 1,
  with a gap of many decades around the library's relative cut of 1e-12
  (LAPACK's SVD, an algorithm independent of the library's pivoted
  Gram-Schmidt);
* the reorganisation itself is exact: the tetrapyd sum over the compressed
  basis, fully symmetrised and expanded through the coefficients Ct (U = P Ct),
  then summed over the 6 slot matchings, gives the oracle's direct Gram.
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

import pyoracle as O
from modalcmb import workloads as W


def pair_products(q):
    R = q.shape[1]
    return np.stack([q[:, x] * q[:, y] for x in range(R) for y in range(x, R)], axis=1)


@pytest.mark.parametrize("name", ["lmax500_r10", "lmax2000_r20", "lmax2500_r30", "gl200_r20", "gl300_r30_batch256"])
def test_pair_products_rank_gap(name):
    cfg = W.WORKLOADS[name]
    R = cfg["R"]
    if "L" in cfg:
        q = W.lspace(cfg["L"], R, cfg["modulated"])["q"]
    else:
        # the k-space basis on the workload's Gauss-Legendre nodes (oracle rule)
        x, _ = O.gauss_legendre(cfg["N"])
        kn = 0.5 * (W.KMAX - W.KMIN) * x + 0.5 * (W.KMAX + W.KMIN)
        q = W.legendre_table(kn / W.KMAX, R - 1)
    s = np.linalg.svd(pair_products(q), compute_uv=False)
    # 2R - 1 directions well above the cut, the rest at rounding level
    assert s[2 * R - 2] / s[0] > 1e-4
    assert s[2 * R - 1] / s[0] < 1e-14


def test_random_basis_pair_products_full_rank():
    # no structure: every pair product is a direction of its own, so the
    # compression (which needs rank <= NP / 2) does not apply
    q = np.random.default_rng(3).standard_normal((121, 6))
    s = np.linalg.svd(pair_products(q), compute_uv=False)
    assert s[-1] / s[0] > 1e-6


def _pair_index(x, y, R):
    x, y = min(x, y), max(x, y)
    return x * R - x * (x - 1) // 2 + (y - x)


@pytest.mark.parametrize("L,R,mod", [(40, 4, True), (33, 5, False)])
def test_compressed_expansion_reproduces_direct_gram(L, R, mod):
    ws = W.lspace(L, R, mod)
    q, v, order = ws["q"], ws["v"], ws["order"]
    U = pair_products(q)                               # [l][pair]
    NP = U.shape[1]
    # U = P Ct with P orthonormal: pivoted QR, rank from the diagonal
    Qf, Rf, piv = scipy.linalg.qr(U, mode="economic", pivoting=True)
    d = np.abs(np.diag(Rf))
    r = int((d > 1e-12 * d[0]).sum())
    assert r == 2 * R - 1
    P = Qf[:, :r]
    Ct = np.zeros((r, NP))
    Ct[:, piv] = Rf[:r, :]
    assert np.abs(P @ Ct - U).max() < 1e-13 * np.abs(U).max()
    # H over the compressed basis from the ordered tetrapyd (oracle weights),
    # fully symmetrised over the three slots
    pts = O.walk_XXX(2, L)
    lib = O.lib()
    w = np.array([lib.or_permsix(int(a), int(b), int(c)) * lib.or_h2_exact(int(a), int(b), int(c))
                  for a, b, c in pts]) * v[pts[:, 0]] * v[pts[:, 1]] * v[pts[:, 2]]
    H = np.einsum("t,ta,tb,tc->abc", w, P[pts[:, 0]], P[pts[:, 1]], P[pts[:, 2]])
    Hs = sum(np.transpose(H, p) for p in itertools.permutations(range(3)))
    # back to pairs: Gs(x, y, z) = sum Ct[a][x] Ct[b][y] Ct[c][z] Hs[a][b][c]
    Gs = np.einsum("ax,by,cz,abc->xyz", Ct, Ct, Ct, Hs)
    # Gamma_mn = (1/36) sum over the 6 matchings of the slots of m and n
    n = order.shape[0]
    G = np.zeros((n, n))
    for m in range(n):
        for k in range(n):
            acc = 0.0
            for s in itertools.permutations(range(3)):
                acc += Gs[_pair_index(order[m][0], order[k][s[0]], R), _pair_index(order[m][1], order[k][s[1]], R),
                          _pair_index(order[m][2], order[k][s[2]], R)]
            G[m, k] = acc / 36.0
    ref = O.gamma_tet(2, L, q, v, order)
    assert np.abs(G - ref).max() < 1e-12 * np.abs(ref).max()
