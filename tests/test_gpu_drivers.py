"""GPU: the command-line drivers end to end on files: Gram -> orthonormalise
-> projection and beta -> reconstruction, against the library calls."""
import os
import subprocess

import numpy as np
import pytest

import modalcmb
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "modal-cmb_amd", "bin")


def run(*args):
    r = subprocess.run([os.path.join(BIN, args[0])] + [str(a) for a in args[1:]], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, (args, r.stdout, r.stderr)
    return r.stdout


def write_order(path, order):
    with open(path, "w") as f:
        f.write("# r s t\n")
        for r, s, t in order.tolist():
            f.write(f"{r} {s} {t}\n")


def test_gamma_and_ortho_drivers(ctx, tmp_path):
    L, R = 80, 4
    order = W.order_lex(R)
    of = str(tmp_path / "order.txt")
    write_order(of, order)
    gf = str(tmp_path / "gamma_TTT")
    out = run("mcmb_gamma", 2, L, of, "modulated", gf, "fp64")
    assert "modes=20" in out
    n = order.shape[0]
    G = modalcmb.read_raw(gf, (n, n))
    ws = W.lspace(L, R, True)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    G0 = ctx.gamma_tet(modalcmb.FP64)
    d = np.sqrt(np.outer(np.diag(G0), np.diag(G0)))
    assert (np.abs(G - G0) / d).max() < 1e-12
    run("mcmb_gamma", 2, L, of, "modulated", gf + "_gl", "gl")
    assert (np.abs(modalcmb.read_raw(gf + "_gl", (n, n)) - G0) / d).max() < 1e-10
    # late-side orthonormalisation files
    run("mcmb_ortho", gf, n, "late", str(tmp_path / "orthol"), str(tmp_path / "lambdal"))
    r, lam, linv, norm, ev = ctx.orthonormalise_late(G, 1)
    assert np.array_equal(modalcmb.read_raw(str(tmp_path / "lambdal"), (n, n)), lam)
    assert np.array_equal(modalcmb.read_raw(str(tmp_path / "orthol"), (n, n)), linv)


def test_eigenmode_and_primordial_drivers(ctx, tmp_path):
    N, R = 48, 3
    order = W.order_lex(R)
    of = str(tmp_path / "order.txt")
    write_order(of, order)
    ef = str(tmp_path / "eigen_4_150")
    out = run("mcmb_eigenmode", N, of, 150.0, 11.14, "-", ef)
    n = order.shape[0]
    beta = modalcmb.read_raw(ef, (n,))
    ks = W.kspace(N, R)
    ctx.kspace_set(ks["kn"], ks["kw"], ks["q"], ks["order"])
    alpha = ctx.project_fwd(np.array([[150.0, 2 * 3.1415927 * 11.14e-2]]), modalcmb.FP64, derivatives=False)[0]
    G = ctx.kgram(modalcmb.FP64)
    if "Cholesky" in out:
        assert np.abs(G @ beta - alpha).max() < 1e-8 * np.abs(alpha).max()
    else:
        # minimum-norm solution on the kept directions: G beta reproduces alpha there
        r, lam, linv, norm, ev = ctx.orthonormalise_late(G, 0)
        assert np.abs(lam[:r] @ (G @ beta - alpha)).max() < 1e-6 * np.abs(lam[:r] @ alpha).max()
    rec = run("mcmb_primordial", of, 150.0, 11.14, ef)
    assert "cosine over the grid points" in rec
    with pytest.raises(AssertionError):
        run("mcmb_ortho", str(tmp_path / "missing"), n, "prim", str(tmp_path / "o"), str(tmp_path / "l"))
