"""GPU parity of the tetrapyd Gram (HIP, through the C ABI) against the oracle."""
import json
import os

import mpmath as mp
import numpy as np
import pytest

import modalcmb
import pyoracle as O
from conftest import GOLDEN
from modalcmb import workloads as W

pytestmark = pytest.mark.gpu


def rel_elem(A, B):
    d = np.sqrt(np.outer(np.abs(np.diag(B)), np.abs(np.diag(B))))
    return float((np.abs(A - B) / d).max())


@pytest.mark.parametrize("lmin,lmax", [(2, 10), (2, 11), (0, 9), (3, 17), (2, 200), (5, 333), (2, 500)])
def test_device_enumeration_bit_exact(ctx, lmin, lmax):
    ws = W.lspace(lmax, 2, lmin=lmin)
    ctx.lspace_set(lmin, lmax, ws["q"], ws["v"], ws["order"])
    pts = ctx.points()
    ref = O.walk_XXX(lmin, lmax)
    assert ctx.npoints() == len(ref)
    assert np.array_equal(pts, ref)


def test_device_count_known_answer(ctx):
    ws = W.lspace(2000, 1)
    ctx.lspace_set(2, 2000, ws["q"], ws["v"], ws["order"])
    assert ctx.npoints() == 334831501          # process.c:65
    head = ctx.points(cap=100_000)
    assert np.array_equal(head, O.walk_XXX_window(2, 2000, 0, 100_000))


@pytest.mark.parametrize("h2", [modalcmb.H2_EXACT, modalcmb.H2_ASYMPTOTIC])
def test_device_weights(ctx, h2):
    L = 300
    ws = W.lspace(L, 2, modulated=True)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"], h2_mode=h2)
    w = ctx.weights()
    pts = O.walk_XXX(2, L)
    f = O.lib().or_h2_exact if h2 == modalcmb.H2_EXACT else O.lib().or_geometric_asym
    v = ws["v"]
    ref = np.array([O.lib().or_permsix(*p) * f(*p) for p in pts.tolist()]) * v[pts[:, 0]] * v[pts[:, 1]] * v[pts[:, 2]]
    assert np.abs(w / ref - 1).max() < 5e-15


def test_gamma_fp64_vs_40_digit_fixture(ctx):
    g = json.load(open(os.path.join(GOLDEN, "gamma_tiny.json")))
    ws = W.lspace(g["L"], g["R"])
    ctx.lspace_set(2, g["L"], ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.FP64)
    n = G.shape[0]
    ref = np.zeros((n, n))
    for a, row in enumerate(g["gamma_upper"]):
        for off, s in enumerate(row):
            ref[a, a + off] = ref[a + off, a] = float(mp.mpf(s))
    assert rel_elem(G, ref) < 1e-13
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("L,R,mod,h2", [(60, 4, False, 0), (97, 5, True, 0), (150, 6, True, 1), (131, 7, False, 0)])
def test_gamma_fp64_vs_oracle_direct(ctx, L, R, mod, h2):
    ws = W.lspace(L, R, mod)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"], h2_mode=h2)
    G = ctx.gamma_tet(modalcmb.FP64)
    ref = O.gamma_tet(2, L, ws["q"], ws["v"], ws["order"], geom=h2)
    assert rel_elem(G, ref) < 1e-12


def test_gamma_fp64_lmax500_vs_oracle(ctx):
    c = W.WORKLOADS["lmax500_r10"]
    ws = W.lspace(c["L"], c["R"], c["modulated"])
    ctx.lspace_set(2, c["L"], ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.FP64)
    ref = O.gamma_gl(2, c["L"], ws["q"], ws["v"], ws["order"])
    assert G.shape == (220, 220)
    assert rel_elem(G, ref) < 1e-10
    assert np.array_equal(G, G.T)
    # deterministic: a second run is bit-identical
    assert np.array_equal(G, ctx.gamma_tet(modalcmb.FP64))


def test_legacy_orthol_3d(ctx):
    # calculate_orthol_3D reads the cached FP64 tetrapyd Gram of the bound
    # context (gamma_3d_host is the cross-Gamma: tests/test_gpu_cross.py)
    import ctypes as C
    lib = modalcmb.load_library()
    ws = W.lspace(80, 4)
    ctx.lspace_set(2, 80, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.FP64)
    lib.mcmb_legacy_bind(ctx._h)
    try:
        res = (C.c_double * 4)()
        assert lib.calculate_orthol_3D(3, 7, res) == 0 and res[0] == G[3, 7]
        assert lib.calculate_orthol_3D(3, 700, res) == -1
    finally:
        lib.mcmb_legacy_bind(None)


@pytest.mark.parametrize("prec", [modalcmb.FP64, modalcmb.MIXED])
def test_slices_sum_to_whole(ctx, prec):
    ws = W.lspace(150, 5, True)
    ctx.lspace_set(2, 150, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(prec)
    parts = [ctx.gamma_tet_slice(prec, r, 5) for r in range(5)]
    assert sum(f for _, f in parts) == pytest.approx(1.0, abs=1e-12)
    S = sum(p for p, _ in parts)
    assert rel_elem(S, G) < (1e-13 if prec == modalcmb.FP64 else 2e-5)


@pytest.mark.parametrize("L,R,mod", [(60, 4, False), (150, 6, True), (500, 10, False)])
def test_gamma_gl_gpu_vs_oracle(ctx, L, R, mod):
    ws = W.lspace(L, R, mod)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_gl()
    ref = O.gamma_gl(2, L, ws["q"], ws["v"], ws["order"])
    assert rel_elem(G, ref) < 1e-12
    assert np.array_equal(G, G.T)


@pytest.mark.parametrize("L,R,mod", [(300, 12, True), (2500, 30, False)])
def test_gamma_gl_compressed_equals_direct_node_sum(ctx, L, R, mod):
    """Legendre bases: the GL Gram runs over the compressed pair basis (rank
    2R - 1); the direct node sum (a permanent per mode pair and node,
    MCMB_GL_PATH=direct) gives the same Gram to rounding."""
    ws = W.lspace(L, R, mod)
    ctx.lspace_set(2, L, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_gl()
    assert ctx.last_basis_rank() == 2 * R - 1
    os.environ["MCMB_GL_PATH"] = "direct"
    try:
        Gd = ctx.gamma_gl()
    finally:
        del os.environ["MCMB_GL_PATH"]
    assert rel_elem(G, Gd) < 1e-13
    assert np.array_equal(G, G.T)
    assert np.array_equal(G, ctx.gamma_gl())


def test_gamma_gl_full_rank_basis_takes_direct_node_sum(ctx):
    """A random basis has full-rank pair products: no compressed basis, and
    the GL Gram is the direct node sum bit for bit."""
    rng = np.random.default_rng(17)
    L, R = 120, 8
    q = rng.standard_normal((L + 1, R))
    v = 1.0 / np.sqrt(1.0 + np.arange(L + 1))
    order = W.order_lex(R)
    ctx.lspace_set(2, L, q, v, order)
    G = ctx.gamma_gl()
    assert ctx.last_basis_rank() == 0
    os.environ["MCMB_GL_PATH"] = "direct"
    try:
        Gd = ctx.gamma_gl()
    finally:
        del os.environ["MCMB_GL_PATH"]
    assert np.array_equal(G, Gd)
    ref = O.gamma_gl(2, L, q, v, order)
    assert rel_elem(G, ref) < 1e-12


def test_gamma_gl_equals_tetrapyd_on_gpu(ctx):
    # two independent GPU evaluations of the same sum at lmax 500
    ws = W.lspace(500, 10, True)
    ctx.lspace_set(2, 500, ws["q"], ws["v"], ws["order"])
    assert rel_elem(ctx.gamma_tet(modalcmb.FP64), ctx.gamma_gl()) < 1e-11


def test_gamma_gl_rejects_inexact_node_count(ctx):
    ws = W.lspace(100, 3)
    ctx.lspace_set(2, 100, ws["q"], ws["v"], ws["order"])
    with pytest.raises(modalcmb.ModalError):
        ctx.gamma_gl(npt=100)


_SEP_VARIANTS = {
    "separable-valu": [("MCMB_SEP_L3", "valu")],
    "separable-staged": [("MCMB_SEP_WEIGHTS", "staged")],
    "separable-pairs": [("MCMB_SEP_COMPRESS", "0")],
    "separable-pairs-valu": [("MCMB_SEP_COMPRESS", "0"), ("MCMB_SEP_L3", "valu")],
    "separable-pairs-staged": [("MCMB_SEP_COMPRESS", "0"), ("MCMB_SEP_WEIGHTS", "staged")],
}


def _gram_paths(ctx, prec, paths=("separable", "factored", "direct")):
    """The same Gram through each contraction order named in paths
    ("separable-valu": the separable order with its l3 contraction on the
    VALU instead of FP64 MFMA; "separable-staged": with the l-space weights
    staged in HBM by their own kernel instead of formed by the l3
    contraction; "separable-pairs...": the same over the pair products
    themselves instead of the compressed basis of their span)."""
    out = {}
    keys = ("MCMB_GRAM_PATH", "MCMB_SEP_L3", "MCMB_SEP_WEIGHTS", "MCMB_SEP_COMPRESS")
    old = {k: os.environ.get(k) for k in keys}
    try:
        for p in paths:
            os.environ["MCMB_GRAM_PATH"] = p.split("-")[0]
            for k in keys[1:]:
                os.environ.pop(k, None)
            for k, val in _SEP_VARIANTS.get(p, []):
                os.environ[k] = val
            out[p] = ctx.gamma_tet(prec)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return out


@pytest.mark.parametrize("lmin,L,R,mod,h2", [(2, 500, 10, True, 0), (3, 401, 12, False, 1), (0, 257, 30, True, 0),
                                             (2, 161, 32, True, 0)])
def test_gamma_contraction_orders_agree(ctx, lmin, L, R, mod, h2):
    # three summation orders of the same FP64 sum: point by point; row moments
    # times pair products; successive contractions over l3, l2, l1 (with the
    # l3 contraction on FP64 MFMA or on the VALU, the weights formed in it or
    # staged), over the compressed basis (the default: Legendre pair products
    # have rank 2R - 1) or over the pairs.  The l3 contraction sums each row
    # in l3 order with fused multiply-adds on every variant, so the separable
    # variants over one basis agree bit for bit.
    ws = W.lspace(L, R, mod, lmin=lmin)
    ctx.lspace_set(lmin, L, ws["q"], ws["v"], ws["order"], h2_mode=h2)
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "separable-valu", "separable-staged", "separable-pairs",
                                         "separable-pairs-valu", "separable-pairs-staged", "factored", "direct"))
    assert ctx.last_basis_rank() == -1   # the direct order ran last
    assert rel_elem(G["separable"], G["direct"]) < 1e-12
    assert rel_elem(G["separable-pairs"], G["direct"]) < 1e-12
    assert rel_elem(G["separable"], G["separable-pairs"]) < 1e-13
    assert rel_elem(G["factored"], G["direct"]) < 1e-12
    assert np.array_equal(G.pop("separable-valu"), G["separable"])
    assert np.array_equal(G.pop("separable-staged"), G["separable"])
    assert np.array_equal(G.pop("separable-pairs-valu"), G["separable-pairs"])
    assert np.array_equal(G.pop("separable-pairs-staged"), G["separable-pairs"])
    for g in G.values():
        assert np.array_equal(g, g.T)
    # the default is the separable order over the compressed basis,
    # bit-identical on repeat
    assert np.array_equal(ctx.gamma_tet(modalcmb.FP64), G["separable"])
    assert ctx.last_basis_rank() == 2 * R - 1


def test_gamma_separable_permuted_mode_order(ctx):
    # modes listed in any order, with their indices in any arrangement
    ws = W.lspace(300, 8, True)
    order = ws["order"].copy()
    rng = np.random.default_rng(7)
    order = order[rng.permutation(len(order))]
    for m in range(len(order)):
        order[m] = order[m][rng.permutation(3)]
    ctx.lspace_set(2, 300, ws["q"], ws["v"], order)
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "direct"))
    assert rel_elem(G["separable"], G["direct"]) < 1e-12


@pytest.mark.parametrize("nranks", [7, 400])
def test_gamma_separable_slices(ctx, nranks):
    # shares that cut the rows of one l1 apart
    ws = W.lspace(300, 9)
    ctx.lspace_set(2, 300, ws["q"], ws["v"], ws["order"])
    G = ctx.gamma_tet(modalcmb.FP64)
    S = np.zeros_like(G)
    for r in range(nranks):
        S += ctx.gamma_tet_slice(modalcmb.FP64, r, nranks)[0]
    assert rel_elem(S, G) < 1e-13
    ref = O.gamma_tet(2, 300, ws["q"], ws["v"], ws["order"])
    assert rel_elem(G, ref) < 1e-12


def test_separable_phases_and_repeat_calls(ctx):
    # per-phase event timings cover the whole call; the cached tables are
    # rebuilt when the setup or the rows change
    ws = W.lspace(300, 10, True)
    ctx.lspace_set(2, 300, ws["q"], ws["v"], ws["order"])
    G1 = ctx.gamma_tet(modalcmb.FP64)
    ph = ctx.last_phases()
    names = [p for p, _ in ph]
    # the first call on the setup factorises the 55 pair products (rank 19);
    # l-space weights are formed inside the l3 contraction (one wave)
    body = ["l3 contraction (row moments)", "l2 contraction (FP64 MFMA)", "symmetrisation",
            "l1 contraction (FP64 MFMA)"]
    assert names == ["basis compression"] + body + ["expansion to pairs", "assembly"]
    assert ctx.last_basis_rank() == 19
    total, _ = ctx.last_kernel_ms()
    assert sum(t for _, t in ph) == pytest.approx(total, rel=0.05, abs=0.05)
    part, _ = ctx.gamma_tet_slice(modalcmb.FP64, 1, 3)
    assert np.array_equal(ctx.gamma_tet(modalcmb.FP64), G1)
    assert [p for p, _ in ctx.last_phases()] == body + ["expansion to pairs", "assembly"]
    # without the per-phase events: the same bits, no phases, the call still timed
    os.environ["MCMB_PHASES"] = "0"
    try:
        assert np.array_equal(ctx.gamma_tet(modalcmb.FP64), G1)
        assert ctx.last_phases() == []
        assert ctx.last_kernel_ms()[0] > 0
    finally:
        del os.environ["MCMB_PHASES"]
    os.environ["MCMB_SEP_COMPRESS"] = "0"
    try:
        Gp = ctx.gamma_tet(modalcmb.FP64)
        assert [p for p, _ in ctx.last_phases()] == ["pair products"] + body + ["assembly"]
        assert ctx.last_basis_rank() == 0
    finally:
        del os.environ["MCMB_SEP_COMPRESS"]
    assert rel_elem(Gp, G1) < 1e-13
    ws2 = W.lspace(300, 10, False)
    ctx.lspace_set(2, 300, ws2["q"], ws2["v"], ws2["order"])
    G2 = ctx.gamma_tet(modalcmb.FP64)
    ref = O.gamma_tet(2, 300, ws2["q"], ws2["v"], ws2["order"])
    assert rel_elem(G2, ref) < 1e-12
    # a Gram through another contraction order records no phases
    os.environ["MCMB_GRAM_PATH"] = "factored"
    try:
        ctx.gamma_tet(modalcmb.FP64)
    finally:
        del os.environ["MCMB_GRAM_PATH"]
    assert ctx.last_phases() == []


@pytest.mark.parametrize("R", [20, 24])
def test_separable_phases_fused_weights(ctx, R):
    # 210 and 300 pairs: the l3 contraction (2 and 4 waves) forms the l-space
    # weights itself, bit-identical to the staged weights; staged, they have
    # a phase of their own
    # (over the pairs themselves: the compressed basis has 2R - 1 columns)
    ws = W.lspace(160, R, True)
    ctx.lspace_set(2, 160, ws["q"], ws["v"], ws["order"])
    G = _gram_paths(ctx, modalcmb.FP64, ("separable-pairs", "separable-pairs-staged", "factored"))
    os.environ["MCMB_SEP_COMPRESS"] = "0"
    try:
        ctx.gamma_tet(modalcmb.FP64)
        names = [p for p, _ in ctx.last_phases()]
        assert names == ["pair products", "l3 contraction (row moments)", "l2 contraction (FP64 MFMA)",
                         "symmetrisation", "l1 contraction (FP64 MFMA)", "assembly"]
        os.environ["MCMB_SEP_WEIGHTS"] = "staged"
        ctx.gamma_tet(modalcmb.FP64)
        assert [p for p, _ in ctx.last_phases()][:2] == ["pair products", "weights"]
    finally:
        os.environ.pop("MCMB_SEP_WEIGHTS", None)
        del os.environ["MCMB_SEP_COMPRESS"]
    assert np.array_equal(G["separable-pairs"], G["separable-pairs-staged"])
    assert rel_elem(G["separable-pairs"], G["factored"]) < 1e-12


@pytest.mark.parametrize("R,compress,h2", [(12, "1", modalcmb.H2_EXACT), (12, "0", modalcmb.H2_EXACT),
                                            (12, "1", modalcmb.H2_ASYMPTOTIC), (20, "0", modalcmb.H2_EXACT)])
def test_separable_row_groups_by_l2(ctx, R, compress, h2):
    # lmax 300 has 299 l1 runs, so the l3 contraction groups rows by l2 (its
    # fused kernel with the legs' roles swapped); its weights are the staged
    # ones bit for bit, and the Gram is that of groups sharing l1 to rounding
    ws = W.lspace(300, R, True)
    ctx.lspace_set(2, 300, ws["q"], ws["v"], ws["order"], h2_mode=h2)
    os.environ["MCMB_SEP_COMPRESS"] = compress
    try:
        G = ctx.gamma_tet(modalcmb.FP64)
        os.environ["MCMB_SEP_WEIGHTS"] = "staged"
        Gs = ctx.gamma_tet(modalcmb.FP64)
        del os.environ["MCMB_SEP_WEIGHTS"]
        os.environ["MCMB_SEP_GROUPS"] = "l1"
        G1 = ctx.gamma_tet(modalcmb.FP64)
    finally:
        for k in ("MCMB_SEP_COMPRESS", "MCMB_SEP_WEIGHTS", "MCMB_SEP_GROUPS"):
            os.environ.pop(k, None)
    assert np.array_equal(G, Gs)
    assert rel_elem(G, G1) < 1e-13
    assert np.array_equal(G, G.T)
    if R == 12 and compress == "1":
        geom = 1 if h2 == modalcmb.H2_ASYMPTOTIC else 0
        assert rel_elem(G, O.gamma_tet(2, 300, ws["q"], ws["v"], ws["order"], geom=geom)) < 1e-12


@pytest.mark.parametrize("threads,block_rows", [("3", None), ("16", None), ("7", "3000")])
def test_separable_tables_any_thread_count(ctx, threads, block_rows):
    # the row group tables are built by host threads over runs of rows and
    # concatenated in run order: any thread count gives the same tables, so
    # the same Gram bit for bit (one block, and blocks of 3000 rows)
    ws = W.lspace(300, 12, True)
    env = {"MCMB_HOST_THREADS": "1", "MCMB_SEP_TABLES": "host"}
    if block_rows:
        env["MCMB_SEP_BLOCK_ROWS"] = block_rows
    os.environ.update(env)
    try:
        ctx.lspace_set(2, 300, ws["q"], ws["v"], ws["order"])
        G1 = ctx.gamma_tet(modalcmb.FP64)
        os.environ["MCMB_HOST_THREADS"] = threads
        ctx.lspace_set(2, 300, ws["q"], ws["v"], ws["order"])      # a new setup: tables rebuilt
        Gt = ctx.gamma_tet(modalcmb.FP64)
    finally:
        for k in env:
            os.environ.pop(k, None)
    assert np.array_equal(G1, Gt)


@pytest.mark.parametrize("R,compress", [(10, "1"), (24, "1"), (10, "0"), (24, "0")])
def test_separable_multi_block(ctx, R, compress):
    # rows cut into blocks of at most 3000 (of 21,841) by MCMB_SEP_BLOCK_ROWS:
    # per-block weights, contractions and symmetrisation, stage 3 accumulating
    # over blocks; against the one-block Gram and the oracle's direct sum;
    # over the compressed basis and over the pairs
    ws = W.lspace(210, R, True)
    ctx.lspace_set(2, 210, ws["q"], ws["v"], ws["order"])
    os.environ["MCMB_SEP_COMPRESS"] = compress
    try:
        G1 = ctx.gamma_tet(modalcmb.FP64)
        os.environ["MCMB_SEP_BLOCK_ROWS"] = "3000"
        Gb = ctx.gamma_tet(modalcmb.FP64)
        names = [p for p, _ in ctx.last_phases()]
        assert ctx.last_basis_rank() == (2 * R - 1 if compress == "1" else 0)
    finally:
        os.environ.pop("MCMB_SEP_BLOCK_ROWS", None)
        del os.environ["MCMB_SEP_COMPRESS"]
    assert "l3, l2, l1 contractions (all blocks)" in names
    assert rel_elem(Gb, G1) < 1e-13
    assert np.array_equal(Gb, Gb.T)
    if R == 10:
        assert rel_elem(Gb, O.gamma_tet(2, 210, ws["q"], ws["v"], ws["order"])) < 1e-12



# R = 15 (120 pairs) and 18 (171 pairs) take the one-wave l3 kernel of 8
# column tiles and the two-wave kernel of 6 tiles (300 of their modes; more
# multipoles than pairs, so the random pair products have full rank)
@pytest.mark.parametrize("R,L", [(6, 120), (12, 120), (15, 130), (18, 180)])
def test_separable_full_rank_basis_contracts_pairs(ctx, R, L):
    # a basis of random values per l: its pair products have full rank, so no
    # compression pays and the pairs are contracted as they are
    ws = W.lspace(L, R, True)
    q = np.random.default_rng(R).standard_normal((L + 1, R))
    if len(ws["order"]) > 300:
        ws["order"] = ws["order"][np.random.default_rng(1).choice(len(ws["order"]), 300, replace=False)]
    ctx.lspace_set(2, L, q, ws["v"], ws["order"])
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "direct"))
    ctx.gamma_tet(modalcmb.FP64)
    assert ctx.last_basis_rank() == 0
    assert "expansion to pairs" not in [p for p, _ in ctx.last_phases()]
    assert rel_elem(G["separable"], G["direct"]) < 1e-12
    assert rel_elem(G["separable"], O.gamma_tet(2, L, q, ws["v"], ws["order"])) < 1e-12


def test_separable_low_rank_non_polynomial_basis(ctx):
    # 8 basis functions spanning only 4 directions (q_4..q_7 are fixed
    # combinations of q_0..q_3, none of them polynomials): the pair products
    # have rank 10 of 36; the compressed Gram matches the pairs' and the
    # oracle's direct sum
    L = 150
    ws = W.lspace(L, 8, True)
    rng = np.random.default_rng(11)
    base = rng.standard_normal((L + 1, 4))
    q = np.concatenate([base, base @ rng.standard_normal((4, 4))], axis=1)
    ctx.lspace_set(2, L, q, ws["v"], ws["order"])
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "separable-pairs"))
    ctx.gamma_tet(modalcmb.FP64)
    assert ctx.last_basis_rank() == 10
    assert rel_elem(G["separable"], G["separable-pairs"]) < 1e-12
    assert rel_elem(G["separable"], O.gamma_tet(2, L, q, ws["v"], ws["order"])) < 1e-12


def test_separable_compressed_wide_basis(ctx):
    # 32 basis functions spanning 12 directions: their 528 pair products have
    # rank 78 > 64, so the compressed contractions take the 160 x 96 tiles
    # and the l3 kernel of 6 column tiles; 400 of the modes
    L, R = 100, 32
    ws = W.lspace(L, R, True)
    rng = np.random.default_rng(34)
    q = rng.standard_normal((L + 1, 12)) @ rng.standard_normal((12, R))
    order = ws["order"][rng.choice(len(ws["order"]), 400, replace=False)]
    ctx.lspace_set(2, L, q, ws["v"], order)
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "separable-pairs"))
    ctx.gamma_tet(modalcmb.FP64)
    assert ctx.last_basis_rank() == 78
    assert rel_elem(G["separable"], G["separable-pairs"]) < 1e-12
    assert rel_elem(G["separable"], O.gamma_tet(2, L, q, ws["v"], order)) < 1e-12
    assert np.array_equal(G["separable"], G["separable"].T)


def _order9_basis(L, p=10, Tmin=2):
    """The reference's late basis of order 9 (create_basis_late, fixed_data.c:
    1923-1950) on multipoles Tmin..L: Legendre P_0..P_{p-2} of (l - Tmin)/(L -
    Tmin) and the columns w/z, w z with x = l(l+1)/(L(L+1)), y = (2l+1)/(2L+1),
    z = sqrt(x), w = y^(1/6); rows indexed by multipole (zero below Tmin)."""
    ls = np.arange(L + 1.0)
    q = np.zeros((L + 1, p + 1))
    tv = ls[Tmin:]
    q[Tmin:, :p - 1] = W.legendre_table((tv - Tmin) / (L - Tmin), p - 2)
    x = tv * (tv + 1) / (L * (L + 1.0))
    y = (2 * tv + 1) / (2 * L + 1.0)
    q[Tmin:, p - 1] = y ** (1 / 6) / np.sqrt(x)
    q[Tmin:, p] = y ** (1 / 6) * np.sqrt(x)
    return q


@pytest.mark.parametrize("kind,rank", [("order9", 0), ("exp", 0), ("cos", 15)])
def test_separable_rank_needs_a_gap(ctx, kind, rank):
    # the pair products are cut to a compressed basis only across a gap of
    # 1e6 in the pivoted Gram-Schmidt residual norms: the reference's order-9
    # late basis and decaying exponentials have residuals that fall through
    # 1e-12 gradually (pivot ratio ~10 at that level), so their pairs are
    # contracted as they are; products of cosines span exactly 15 of 36
    # directions and are compressed.  Every result matches the pairs and the
    # oracle's direct sum.
    L = 200
    ls = np.arange(L + 1.0)
    if kind == "order9":
        q = _order9_basis(L)
    elif kind == "exp":
        q = np.stack([np.exp(-r * ls / L) for r in range(8)], 1)
    else:
        q = np.stack([np.cos((r + 0.5) * np.pi * ls / L * 0.7) for r in range(8)], 1)
    R = q.shape[1]
    ws = W.lspace(L, R, True)
    ctx.lspace_set(2, L, q, ws["v"], ws["order"])
    G = _gram_paths(ctx, modalcmb.FP64, ("separable", "separable-pairs"))
    ctx.gamma_tet(modalcmb.FP64)
    assert ctx.last_basis_rank() == rank
    ref = O.gamma_tet(2, L, q, ws["v"], ws["order"])
    assert rel_elem(G["separable"], ref) < 1e-12
    assert rel_elem(G["separable-pairs"], ref) < 1e-12


@pytest.mark.parametrize("lmax,R,block_rows,weights", [(300, 12, None, None), (300, 12, None, "staged"),
                                                       (2500, 30, None, None), (2500, 30, "1500000", None),
                                                       (2500, 30, None, "staged")])
def test_separable_device_row_groups_equal_host(ctx, lmax, R, block_rows, weights):
    # row groups sharing l2 formed on the device (one wave per l2 and parity)
    # are the host walk's groups, l3 ranges and weight offsets: the same Gram
    # bit for bit, with fused weights and with staged ones (which read the
    # offsets), in one block and in several
    ws = W.lspace(lmax, R, True)
    env = {}
    if block_rows:
        env["MCMB_SEP_BLOCK_ROWS"] = block_rows
    if weights:
        env["MCMB_SEP_WEIGHTS"] = weights
    os.environ.update(env)
    try:
        ctx.lspace_set(2, lmax, ws["q"], ws["v"], ws["order"])
        Gd = ctx.gamma_tet(modalcmb.FP64)
        os.environ["MCMB_SEP_TABLES"] = "host"
        ctx.lspace_set(2, lmax, ws["q"], ws["v"], ws["order"])      # a new setup: tables rebuilt on the host
        Gh = ctx.gamma_tet(modalcmb.FP64)
    finally:
        for k in list(env) + ["MCMB_SEP_TABLES"]:
            os.environ.pop(k, None)
    assert np.array_equal(Gd, Gh)
    names = [n for n, _ in ctx.last_setup()]
    assert "row groups and tables (host)" in names
