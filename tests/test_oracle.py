"""Pins the oracle (CPU restatement) against the reference's own known
answers, closed forms, and the reference itself built from its sources
(oracle/_ref, when present).  CPU only."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from hmat_lib_amd import workloads as WL
from oracle import oracle as O
from tests import helpers as H

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _bfs_sizes(t):
    lev = T.levels(t.contents.root, t.contents.height)
    out = [t.contents.root.contents.m]
    for d in range(1, t.contents.height):
        out += [p.contents.m for p in lev[d]]
    for p in lev[-1]:
        out += [p.contents.children[0].leaf.contents.data.diagonal.m,
                p.contents.children[3].leaf.contents.data.diagonal.m]
    return out


def test_block_sizes_halves_and_custom():
    gold = json.load(open(os.path.join(GOLD, "block_sizes.json")))
    lib = O.oracle()
    ierr = C.c_int(0)
    for h, exp in gold["halves"]:
        t = lib.orc_allocate_tree(h, C.byref(ierr))
        lib.orc_block_sizes_halves(t, exp[0])
        assert _bfs_sizes(t) == exp
        # custom with the leaves of the halves split gives the same tree
        leaves = exp[-(1 << h):]
        t2 = lib.orc_allocate_tree(h, C.byref(ierr))
        lib.orc_block_sizes_custom(t2, (C.c_int * len(leaves))(*leaves))
        assert _bfs_sizes(t2) == exp
        lib.orc_free_tree(t)
        lib.orc_free_tree(t2)
    for h, ms, exp in gold["uneven_custom"]:
        t = lib.orc_allocate_tree(h, C.byref(ierr))
        lib.orc_block_sizes_custom(t, (C.c_int * len(ms))(*ms))
        assert _bfs_sizes(t) == exp
        lib.orc_free_tree(t)


def test_planted_mistake_is_caught():
    """A block-size rule with the halves swapped (top-left gets floor) must
    fail the golden table: the table pins the split orientation."""
    gold = json.load(open(os.path.join(GOLD, "block_sizes.json")))
    h, exp = gold["halves"][1]  # m = 9

    def wrong(m, depth):
        if depth == h:
            return [m]
        return wrong(m // 2, depth + 1) + wrong(m - m // 2, depth + 1)

    assert wrong(exp[0], 0) != exp[-(1 << h):]


def test_splitmix64_known_value():
    # published first output of splitmix64 from seed 0
    assert int(WL.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
    u = WL.uniform(123, 100000)
    assert 0.0 <= u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.01


def test_morton_order_is_z_curve():
    pts = np.array([[0.9, 0.9], [0.1, 0.1], [0.9, 0.1], [0.1, 0.9]])
    assert list(WL.morton_order(pts)) == [1, 2, 3, 0]


@pytest.mark.parametrize("case_i", range(7))
def test_compress_block_known_answers(case_i):
    """test_constructors.c:725-873 on the oracle: rank and U V^T exactly;
    factors up to the sign of each (u_j, v_j) pair."""
    case = json.load(open(os.path.join(GOLD, "compress_block.json")))["cases"][case_i]
    mf = case["m_full"]
    full = {"laplacian": H.laplacian, "laplacian_corners": H.laplacian_corners,
            "identity": np.eye, "zeros": lambda m: np.zeros((m, m))}[case["full"]](mf)
    full = np.asfortranarray(full)
    m, n, r0, c0 = case["m"], case["n"], case["r0"], case["c0"]
    blk = np.asfortranarray(full[r0:r0 + m, c0:c0 + n].copy())
    pu, pv = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    s = C.c_int(0)
    ierr = C.c_int(0)
    O.oracle().orc_compress_block(m, n, capi.dptr(blk), m, 1e-8, C.byref(pu), C.byref(pv),
                                  C.byref(s), None, C.byref(ierr))
    assert ierr.value == 0 and s.value == case["s"]
    u = np.ctypeslib.as_array(pu, shape=(m * s.value,)).reshape(m, s.value, order="F")
    v = np.ctypeslib.as_array(pv, shape=(n * s.value,)).reshape(n, s.value, order="F")
    ue = np.zeros((m, s.value))
    ve = np.zeros((n, s.value))
    for k, val in case["u_exp"].items():
        i, j = map(int, k.split(","))
        ue[i, j] = val
    for k, val in case["v_exp"].items():
        i, j = map(int, k.split(","))
        ve[i, j] = val
    np.testing.assert_allclose(u @ v.T, ue @ ve.T, atol=1e-10)
    a, b = H.sign_normalize(u, v)
    c, d = H.sign_normalize(ue, ve)
    np.testing.assert_allclose(a, c, atol=1e-10)
    np.testing.assert_allclose(b, d, atol=1e-10)


@pytest.mark.parametrize("height", [1, 2, 3])
@pytest.mark.parametrize("kind", ["L", "L0.5S", "L0.5A"])
def test_laplacian_products_known_answers(height, kind):
    gold = json.load(open(os.path.join(GOLD, "products.json")))
    m = gold["m"]
    a = {"L": H.laplacian(m), "L0.5S": H.laplacian_corners(m),
         "L0.5A": H.laplacian_corners(m, symmetric=False)}[kind]
    t = H.oracle_compress(a, height, gold["thr"])
    exp = np.zeros(m)
    for k, v in gold["laplacian_vector_10"][kind].items():
        exp[int(k)] = v
    np.testing.assert_allclose(H.oracle_matvec(t, np.full(m, 10.0)), exp, atol=1e-10)
    X = np.full((m, m), 10.0, order="F")
    np.testing.assert_allclose(H.oracle_matmat(t, X, "hd"), np.tile(exp[:, None], (1, m)),
                               atol=1e-10)
    key = kind + ("_transpose" if kind == "L0.5A" else "")
    expt = np.zeros(m)
    for k, v in gold["laplacian_dense_10_rows"][key].items():
        expt[int(k)] = v
    np.testing.assert_allclose(H.oracle_matmat(t, X, "htd"), np.tile(expt[:, None], (1, m)),
                               atol=1e-10)
    H.oracle_free(t)


def test_sorted_exponential_is_rank_one_closed_form():
    """exp(-|x_i-x_j|/l) with x sorted: every off-diagonal block is
    exp(x_i/l) exp(-x_j/l) (or transposed), i.e. exactly rank 1."""
    w = WL.Workload("t", 512, 3, 0, 0.1, 0.0, 1e-10, 4, 1, 99)
    x = w.points()
    a = np.asfortranarray(np.exp(-np.abs(x[:, None] - x[None, :]) / w.ell))
    t = H.oracle_compress(a, 3, 1e-10)
    ft = T.flatten(t)
    for key, (u, v) in ft.blocks.items():
        assert u.shape[1] == 1, key
        (r0, r1), (c0, c1) = ft.block_ranges(*key)
        xi, xj = x[r0:r1], x[c0:c1]
        closed = np.exp(-np.abs(xi[:, None] - xj[None, :]) / w.ell)
        np.testing.assert_allclose(u @ v.T, closed, rtol=0, atol=1e-14)
    H.oracle_free(t)


def test_recompress_golden_on_oracle():
    data = json.load(open(os.path.join(GOLD, "recompress.json")))
    from tests.test_gpu_recompress import _node, _read, _sparse
    for method, fn in (("qr", O.oracle().orc_recompress),
                       ("large_s", O.oracle().orc_recompress_large_s)):
        for case in data[method]:
            m, n, s = case["m"], case["n"], case["s"]
            nd = _node(_sparse((m, s), case["u"]), _sparse((n, s), case["v"]))
            ierr = C.c_int(0)
            fn(C.byref(nd), min(m, n), 1e-8, C.byref(ierr))
            assert ierr.value == 0 and nd.s == case["s_exp"]
            u, v = _read(nd)
            ue = _sparse((m, nd.s), case["u_exp"])
            ve = _sparse((n, nd.s), case["v_exp"])
            np.testing.assert_allclose(u @ v.T, ue @ ve.T, atol=1e-10)
            if nd.s == 3:
                a, b = H.sign_normalize(u, v)
                c, d = H.sign_normalize(ue, ve)
                np.testing.assert_allclose(a, c, atol=1e-10)
                np.testing.assert_allclose(b, d, atol=1e-10)


# ------------------------------------------------ against the reference
needs_ref = pytest.mark.skipif(not O.have_ref(),
                               reason="oracle/_ref not built (reference sources absent)")


def _ref_compress(a, height, thr, ms=None):
    lib = O.ref()
    ierr = C.c_int(0)
    t = lib.allocate_tree_monolithic(height, C.byref(ierr))
    work = np.asfortranarray(a.copy())
    ms_arr = (C.c_int * len(ms))(*ms) if ms is not None else None
    lib.dense_to_tree_hodlr(t, a.shape[0], ms_arr, capi.dptr(work), thr, C.byref(ierr))
    assert ierr.value == 0
    return t


@needs_ref
def test_oracle_matches_reference_bitwise():
    """Same LAPACK, same call sequence: the restatement reproduces the
    reference bit for bit (compression and all products)."""
    from tests.test_gpu_compress import MS155
    mats = H.glibc_decay_matrices(155, 1.0, True, 10, 10)
    idx = 0
    rng = np.random.default_rng(0)
    for kind in ("sorted", "unsorted"):
        for custom in (False, True):
            for height in range(1, 6):
                a = mats[idx][1]
                idx += 1
                ms = MS155[height] if custom else None
                to = H.oracle_compress(a, height, 1e-8, ms)
                tr = _ref_compress(a, height, 1e-8, ms)
                fo, fr = T.flatten(to), T.flatten(tr)
                for key in fo.blocks:
                    np.testing.assert_array_equal(fo.blocks[key][0], fr.blocks[key][0])
                    np.testing.assert_array_equal(fo.blocks[key][1], fr.blocks[key][1])
                X = np.asfortranarray(rng.standard_normal((155, 7)))
                ierr = C.c_int(0)
                rl = O.ref()
                for name, mode in (("multiply_hodlr_dense", "hd"),
                                   ("multiply_hodlr_transpose_dense", "htd")):
                    out = np.zeros((155, 7), order="F")
                    getattr(rl, name)(tr, capi.dptr(X), 7, 155, capi.dptr(out), 155,
                                      C.byref(ierr))
                    np.testing.assert_array_equal(out, H.oracle_matmat(to, X, mode))
                x = rng.standard_normal(155)
                y = np.zeros(155)
                rl.multiply_vector(tr, capi.dptr(x), capi.dptr(y), C.byref(ierr))
                np.testing.assert_array_equal(y, H.oracle_matvec(to, x))
                Xr = np.asfortranarray(rng.standard_normal((3, 155)))
                out = np.zeros((3, 155), order="F")
                rl.multiply_dense_hodlr(tr, capi.dptr(Xr), 3, 3, capi.dptr(out), 3,
                                        C.byref(ierr))
                np.testing.assert_array_equal(out, H.oracle_matmat(to, Xr, "dh"))
                H.oracle_free(to)
                rl.free_tree_hodlr(C.byref(tr))


@needs_ref
def test_oracle_recompress_matches_reference():
    from tests.test_gpu_recompress import _node, _read
    rng = np.random.default_rng(3)
    for _ in range(10):
        m, n = rng.integers(20, 90, size=2)
        for fo, fr, big in ((O.oracle().orc_recompress, O.ref().ref_recompress, False),
                            (O.oracle().orc_recompress_large_s,
                             O.ref().ref_recompress_large_s, True)):
            # QR method needs s < min(m, n); the dense method writes
            # min(m, n) columns into U/V and so needs s >= min(m, n)
            s = int(min(m, n) + 3) if big else int(rng.integers(2, min(m, n) - 1))
            u = rng.standard_normal((m, s)) * np.logspace(0, -14, s)
            v = rng.standard_normal((n, s))
            a, b = _node(u, v), _node(u, v)
            e1, e2 = C.c_int(0), C.c_int(0)
            fo(C.byref(a), int(min(m, n)), 1e-8, C.byref(e1))
            fr(C.byref(b), int(min(m, n)), 1e-8, C.byref(e2))
            assert a.s == b.s
            ua, va = _read(a)
            ub, vb = _read(b)
            np.testing.assert_array_equal(ua, ub)
            np.testing.assert_array_equal(va, vb)


@needs_ref
def test_reference_fixture_laplacian_on_reference():
    """The golden products hold on the reference build itself."""
    gold = json.load(open(os.path.join(GOLD, "products.json")))
    m = gold["m"]
    for kind, a in (("L", H.laplacian(m)), ("L0.5S", H.laplacian_corners(m))):
        t = _ref_compress(a, 2, gold["thr"])
        x = np.full(m, 10.0)
        y = np.zeros(m)
        ierr = C.c_int(0)
        O.ref().multiply_vector(t, capi.dptr(x), capi.dptr(y), C.byref(ierr))
        exp = np.zeros(m)
        for k, v in gold["laplacian_vector_10"][kind].items():
            exp[int(k)] = v
        np.testing.assert_allclose(y, exp, atol=1e-10)
        O.ref().free_tree_hodlr(C.byref(t))


# ------------------------------------------------- rank predictor oracle
def test_mlp_oracle_gradient_finite_differences():
    """The oracle's MLP gradient equals central differences of its loss, for
    every parameter block (a dropped chain-rule term, a transposed weight
    or a wrong tanh derivative fails here)."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(3)
    nin, nh, B = 5, 7, 11
    p = rng.standard_normal(M.nparams(nin, nh)) * 0.5
    X = rng.standard_normal((B, nin))
    t = rng.standard_normal(B)
    loss, g = M.loss_and_grad(p, X, t, nh)
    h = 1e-6
    num = np.empty_like(p)
    for k in range(p.size):
        e = np.zeros_like(p)
        e[k] = h
        num[k] = (M.loss_and_grad(p + e, X, t, nh)[0] - M.loss_and_grad(p - e, X, t, nh)[0]) / (2 * h)
    np.testing.assert_allclose(g, num, rtol=1e-6, atol=1e-9)
    assert loss > 0


def test_mlp_oracle_forward_formula():
    """Forward pass vs a per-sample loop over the formula."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(4)
    nin, nh = 3, 4
    p = rng.standard_normal(M.nparams(nin, nh))
    X = rng.standard_normal((6, nin))
    W1, b1, W2, b2, w3, b3 = M.unpack(p, nin, nh)
    y, _, _ = M.forward(p, X, nh)
    for b in range(6):
        a1 = [np.tanh(sum(X[b, i] * W1[i, j] for i in range(nin)) + b1[j]) for j in range(nh)]
        a2 = [np.tanh(sum(a1[i] * W2[i, j] for i in range(nh)) + b2[j]) for j in range(nh)]
        assert abs(y[b] - (sum(a2[j] * w3[j] for j in range(nh)) + b3)) < 1e-12


def test_rank_features_definition():
    """Features of hand-made spectra: sigma_i / sigma_0 = 10^-i gives
    log10 ratios -i/10, and 8 of the 15 ratios reach thr = 1e-8."""
    from oracle import mlp_oracle as M
    sig = np.array([[10.0 ** (-i) for i in range(16)], [0.0] * 16])
    f = M.rank_features(np.array([[1.0, 2.0], [3.0, 4.0]]), sig, 1e-8)
    assert f.shape == (2, 19)
    np.testing.assert_allclose(f[0, 2:17], -np.arange(1, 16) / 10.0)
    assert f[0, 17] == 8 / 16
    np.testing.assert_allclose(f[:, 18], -0.8)
    np.testing.assert_allclose(f[1, 2:17], -2.0)  # zero block: floored
    assert f[1, 17] == 0


def test_rank_features_library_matches_oracle():
    """The library's host feature function (used by compression) equals the
    oracle's on random spectra."""
    import ctypes as C
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(5)
    nb, ne = 9, 3
    sig = np.sort(10.0 ** rng.uniform(-14, 0, size=(nb, 16)), axis=1)[:, ::-1].copy()
    sig[2] = 0.0
    extra = rng.standard_normal((nb, ne)).astype(np.float32)
    out = np.zeros((nb, ne + 17), dtype=np.float32)
    rc = capi.hip().hm_hip_rank_features(
        nb, ne, extra.ctypes.data_as(C.POINTER(C.c_float)),
        np.ascontiguousarray(sig).ctypes.data_as(C.POINTER(C.c_double)), 1e-8,
        out.ctypes.data_as(C.POINTER(C.c_float)))
    assert rc == 0
    np.testing.assert_allclose(out, M.rank_features(extra, sig, 1e-8), rtol=1e-6, atol=1e-7)


# ------------------------------------------------------- H x H oracle
@pytest.mark.parametrize("height", [1, 2, 3])
def test_hh_oracle_known_answers(height):
    """The reference's H x H known answers (test_hodlr_hodlr_algebra.c:
    100-339, DELTA 1e-10 on the reconstruction) hold on the oracle; the
    non-symmetric corner cases catch a transposed factor or a block taken
    from the wrong side."""
    for name, a, b, exp, same in H.hh_known_answers():
        np.testing.assert_allclose(a @ b, exp, atol=0, err_msg=name)  # fixture transcription
        ta = H.oracle_compress(a, height, 1e-8)
        tb = ta if same else H.oracle_compress(b, height, 1e-8)
        tc = H.oracle_hh(ta, tb, 1e-8)
        np.testing.assert_allclose(T.flatten(tc).to_dense(), exp, atol=1e-10, err_msg=name)
        assert tc.contents.root.contents.m == a.shape[0]
        for t in {id(x): x for x in (ta, tb, tc)}.values():
            H.oracle_free(t)


def test_hh_oracle_exact_algebra_random_trees():
    """Random trees (uneven leaves, ranks 1-7, some blocks with s_tot >=
    min(m, n) so both recompression methods run): with a threshold far
    below the singular values kept, the product tree equals the dense
    product of the two dense matrices (a dropped term, a wrong offset or a
    wrong ancestor fails this)."""
    rng = np.random.default_rng(11)
    lib = capi.hmat()
    for height, leaves in ((2, [9, 14, 6, 11]), (3, [7, 12, 5, 9, 13, 8, 6, 10])):
        keep = T.Keep()
        fa = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(1, 5)))
        fb = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(1, 8)))
        ta, tb = T.fill_tree(lib, fa, keep), T.fill_tree(lib, fb, keep)
        tc = H.oracle_hh(ta, tb, 1e-12)
        exp = fa.to_dense() @ fb.to_dense()
        assert H.rel_err(T.flatten(tc).to_dense(), exp) < 1e-13
        H.oracle_free(tc)
        T.detach_and_free(lib, ta)
        T.detach_and_free(lib, tb)


@needs_ref
def test_hh_oracle_matches_reference_bitwise():
    """multiply_hodlr_hodlr of the reference built from its sources and the
    restatement agree bit for bit (same LAPACK, same call sequence), on the
    glibc decay matrices (m = 155, heights 1-5, halves and custom leaves)
    and on a tree multiplied by itself."""
    from tests.test_gpu_compress import MS155
    mats = H.glibc_decay_matrices(155, 1.0, True, 6, 6)
    rl = O.ref()
    k = 0
    for height in range(1, 6):
        for custom in (False, True):
            a, b = mats[k % 12][1], mats[(k + 5) % 12][1]
            k += 1
            ms = MS155[height] if custom else None
            to_a, to_b = H.oracle_compress(a, height, 1e-8, ms), H.oracle_compress(b, height, 1e-8, ms)
            tr_a, tr_b = _ref_compress(a, height, 1e-8, ms), _ref_compress(b, height, 1e-8, ms)
            for same in (False, True):
                ierr = C.c_int(0)
                tr_c = rl.allocate_tree_monolithic(height, C.byref(ierr))
                info = rl.multiply_hodlr_hodlr(tr_a, tr_a if same else tr_b, tr_c, 1e-8,
                                               C.byref(ierr))
                assert ierr.value == 0 and info == 0
                to_c = H.oracle_hh(to_a, to_a if same else to_b, 1e-8)
                fo, fr = T.flatten(to_c), T.flatten(tr_c)
                for i in range(len(fo.diag)):
                    np.testing.assert_array_equal(fo.diag[i], fr.diag[i])
                for key in fr.blocks:
                    np.testing.assert_array_equal(fo.blocks[key][0], fr.blocks[key][0])
                    np.testing.assert_array_equal(fo.blocks[key][1], fr.blocks[key][1])
                H.oracle_free(to_c)
                rl.free_tree_hodlr(C.byref(tr_c))
            for t in (to_a, to_b):
                H.oracle_free(t)
            for t in (tr_a, tr_b):
                rl.free_tree_hodlr(C.byref(t))
