"""Generate the golden fixtures under tests/golden/ with mpmath (40 digits).

Nothing here reads or runs the reference; the fixtures come from textbook
mathematics the reference's functions must agree with:

* threej.json   -- (l1 l2 l3; 0 0 0)^2 from the factorial formula and
                   h^2 = (2l1+1)(2l2+1)(2l3+1)/(4 pi) (3j)^2, incl. edge cases
* gl.json       -- Gauss-Legendre nodes/weights (n = 5, 20, 64) by Newton on
                   mpmath.legendre from numpy's leggauss start
* legendre.json -- P_r(x), r < 30, at a few x by mpmath.legendre
* gamma_tiny.json -- brute-force Gram over T(2, 24), R = 3 (10 modes), with
                   exact 3j, v(l) = l(l+1)/1000, shifted Legendre basis on l/24,
                   all in 40-digit arithmetic
* walk_counts.json -- brute-force counts of T(lmin, lmax) by triple loop

Run:  python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def threej_sq(l1, l2, l3):
    L = l1 + l2 + l3
    if L % 2 or l1 + l2 < l3 or l2 + l3 < l1 or l3 + l1 < l2:
        return mp.mpf(0)
    g = L // 2
    f = mp.factorial
    return (f(L - 2 * l1) * f(L - 2 * l2) * f(L - 2 * l3) / f(L + 1)) * \
        (f(g) / (f(g - l1) * f(g - l2) * f(g - l3))) ** 2


def h2(l1, l2, l3):
    return (2 * l1 + 1) * (2 * l2 + 1) * (2 * l3 + 1) / (4 * mp.pi) * threej_sq(l1, l2, l3)


def make_threej():
    triples = [(0, 0, 0), (1, 1, 0), (2, 2, 2), (2, 3, 5), (3, 4, 5), (2, 2, 4), (10, 10, 10),
               (5, 7, 12), (100, 200, 300), (250, 250, 500), (499, 500, 501), (2, 499, 499),
               (1000, 1500, 2496), (1234, 1500, 2000), (2000, 2000, 2000), (2, 2499, 2499),
               (2500, 2500, 2500), (1, 2500, 2499), (777, 888, 1555), (3, 4, 6), (2, 3, 4)]
    rng = np.random.default_rng(7)
    while len(triples) < 60:
        a, b = sorted(rng.integers(2, 2501, size=2).tolist())
        lo, hi = b - a, min(a + b, 2500)
        c = int(rng.integers(max(lo, b), hi + 1)) if hi >= max(lo, b) else b
        if (a + b + c) % 2:
            c = c - 1 if c - 1 >= max(lo, b) else c + 1
        if c > 2500 or (a + b + c) % 2 or c < b or c > a + b:
            continue
        triples.append((a, b, c))
    out = [{"l": list(t), "threej_sq": mp.nstr(threej_sq(*t), 35), "h2": mp.nstr(h2(*t), 35)}
           for t in triples]
    return out


def gl_nodes(n):
    x0, _ = np.polynomial.legendre.leggauss(n)
    xs, ws = [], []
    for x in x0:
        r = mp.findroot(lambda t: mp.legendre(n, t), mp.mpf(float(x)))
        d = mp.diff(lambda t: mp.legendre(n, t), r)
        xs.append(r)
        ws.append(2 / ((1 - r * r) * d * d))
    return [mp.nstr(v, 35) for v in xs], [mp.nstr(v, 35) for v in ws]


def make_legendre():
    xs = [-1, -0.999, -0.5, 0.0, 0.123456789, 0.75, 0.9999, 1]
    return [{"x": x, "P": [mp.nstr(mp.legendre(r, mp.mpf(x)), 35) for r in range(30)]} for x in xs]


def make_gamma_tiny(L=24, R=3, lmin=2):
    order = [(r, s, t) for r in range(R) for s in range(r, R) for t in range(s, R)]
    n = len(order)

    def P(r, x):
        return mp.legendre(r, x)

    q = [[P(r, mp.mpf(2 * l) / L - 1) for r in range(R)] for l in range(L + 1)]
    v = [mp.mpf(l * (l + 1)) / 1000 for l in range(L + 1)]
    G = [[mp.mpf(0)] * n for _ in range(n)]
    for i in range(lmin, L + 1):
        for j in range(i, L + 1):
            for k in range(j, min(i + j, L) + 1):
                if (i + j + k) % 2:
                    continue
                mu = 1 if i == j == k else (3 if (i == j or j == k or i == k) else 6)
                w = mu * h2(i, j, k) * v[i] * v[j] * v[k]
                Q = []
                for (r, s, t) in order:
                    Q.append((q[i][r] * q[j][s] * q[k][t] + q[i][s] * q[j][t] * q[k][r]
                              + q[i][t] * q[j][r] * q[k][s] + q[i][t] * q[j][s] * q[k][r]
                              + q[i][s] * q[j][r] * q[k][t] + q[i][r] * q[j][t] * q[k][s]) / 6)
                for a in range(n):
                    wa = w * Q[a]
                    for b in range(a, n):
                        G[a][b] += wa * Q[b]
    return {"L": L, "R": R, "lmin": lmin, "order": order,
            "gamma_upper": [[mp.nstr(G[a][b], 30) for b in range(a, n)] for a in range(n)]}


def make_counts():
    out = []
    for lmin, lmax in [(2, 10), (2, 11), (0, 7), (3, 9), (2, 40), (5, 33)]:
        c = 0
        for i in range(lmin, lmax + 1):
            for j in range(i, lmax + 1):
                for k in range(j, lmax + 1):
                    if k <= i + j and (i + j + k) % 2 == 0:
                        c += 1
        out.append({"lmin": lmin, "lmax": lmax, "count": c})
    return out


def main():
    with open(os.path.join(HERE, "threej.json"), "w") as f:
        json.dump(make_threej(), f, indent=0)
    gl = {}
    for n in (5, 20, 64):
        x, w = gl_nodes(n)
        gl[str(n)] = {"x": x, "w": w}
    with open(os.path.join(HERE, "gl.json"), "w") as f:
        json.dump(gl, f, indent=0)
    with open(os.path.join(HERE, "legendre.json"), "w") as f:
        json.dump(make_legendre(), f, indent=0)
    with open(os.path.join(HERE, "gamma_tiny.json"), "w") as f:
        json.dump(make_gamma_tiny(), f, indent=0)
    with open(os.path.join(HERE, "walk_counts.json"), "w") as f:
        json.dump(make_counts(), f, indent=0)


if __name__ == "__main__":
    main()
