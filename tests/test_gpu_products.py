"""HODLR x vector / HODLR x dense on the GPU vs the oracle restatement and
vs exact dense algebra, through both the device C-ABI and the drop-in API.

Tolerance: 1e-12 relative (FP64 throughout; only the summation order
differs from the reference's BLAS loops)."""
import ctypes as C
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from tests import helpers as H

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mfma_lane_layout(gpu):
    _say("case: mfma selftest")
    out = (C.c_double * 256)()
    capi.check(capi.hip().hm_hip_selftest_mfma(out), "selftest")
    got = np.array(out[:]).reshape(16, 16)
    i = np.arange(16)
    A = 1.0 + i[:, None] + 16.0 * np.arange(4)[None, :]
    B = 1.0 + 2.0 * i[None, :] - np.arange(4)[:, None]
    np.testing.assert_array_equal(got, A @ B)


def test_mfma_bf16_lane_layout(gpu):
    """v_mfma_f32_16x16x32_bf16 lane layout the sliced kernels assume: lane L
    holds A[L%16][8(L/16)+e], B[8(L/16)+e][L%16]; value v of lane L is
    D[4(L/16)+v][L%16].  Integer inputs make every product exact."""
    _say("case: mfma bf16 selftest")
    rng = np.random.default_rng(0)
    A = rng.integers(-8, 9, size=(16, 32)).astype(np.float32)
    B = rng.integers(-8, 9, size=(32, 16)).astype(np.float32)
    raw = np.zeros(256, dtype=np.float32)
    FP = C.POINTER(C.c_float)
    capi.check(capi.hip().hm_hip_selftest_mfma_bf16(A.ctypes.data_as(FP), B.ctypes.data_as(FP),
                                                    raw.ctypes.data_as(FP)), "selftest bf16")
    L = np.arange(64)
    got = np.zeros((16, 16), dtype=np.float32)
    for v in range(4):
        got[4 * (L // 16) + v, L % 16] = raw[4 * L + v]
    np.testing.assert_array_equal(got, A @ B)


@pytest.mark.parametrize("fmt,ks,lim", [(1, 32, 8), (2, 64, 127), (3, 16, 512)])
def test_mfma_lowp_lane_layout(gpu, fmt, ks, lim):
    """One 16 x 16 block of the sliced GEMM per format, from 16-byte lane
    fragments (lane L: A[L%16][e0 + e], B[e0 + e][L%16], e0 = ks/4 (L/16)):
    BF16 one 16x16x32 MFMA, INT8 one v_mfma_i32_16x16x64_i8 (|q| <= 127),
    FP32 four 16x16x4 MFMAs; value v of lane L is D[4(L/16)+v][L%16].
    Integer inputs keep every product and sum exact."""
    _say(f"case: mfma lowp selftest format={fmt}")
    rng = np.random.default_rng(fmt)
    A = rng.integers(-lim, lim + 1, size=(16, ks)).astype(np.float32)
    B = rng.integers(-lim, lim + 1, size=(ks, 16)).astype(np.float32)
    raw = np.zeros(256, dtype=np.float32)
    FP = C.POINTER(C.c_float)
    capi.check(capi.hip().hm_hip_selftest_mfma_lowp(fmt, A.ctypes.data_as(FP),
                                                    B.ctypes.data_as(FP),
                                                    raw.ctypes.data_as(FP)), "selftest lowp")
    L = np.arange(64)
    got = np.zeros((16, 16), dtype=np.float64)
    for v in range(4):
        got[4 * (L // 16) + v, L % 16] = raw[4 * L + v]
    np.testing.assert_array_equal(got, A.astype(np.float64) @ B.astype(np.float64))


CASES = [
    ("uniform h3 s2", [32] * 8, lambda l, j: 2),
    ("uniform h4 mixed", [64] * 16, lambda l, j: 1 + (l + j) % 5),
    ("ragged h3", [10, 11, 1, 9, 5, 20, 7, 6], lambda l, j: 1 + (j % 3)),
    ("ragged zeros h2", [20, 1, 1, 0], lambda l, j: 1),
    ("big leaves h2", [300, 129, 77, 256], lambda l, j: 17 + l),
    ("rank 40 h5", [96] * 32, lambda l, j: 40 - 5 * l),
    # level-1 blocks span 5200 rows: two K segments in the block path
    ("two segments h2", [2600] * 4, lambda l, j: 70 if l == 1 else 30 + j),
]

# W = F_in^T X strategy: chosen by the library ("auto"), or forced to the
# per-chunk partials ("chunk") or the per-block split-K GEMM ("block").
WPATHS = {"auto": None, "chunk": "0", "block": "1"}


@pytest.mark.parametrize("wpath", list(WPATHS))
@pytest.mark.parametrize("name,leaves,ranks", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("R", [1, 3, 64, 100])
def test_products_vs_dense(gpu, monkeypatch, name, leaves, ranks, R, wpath):
    from hmat_lib_amd import device as D
    if WPATHS[wpath] is None:
        monkeypatch.delenv("HM_WPATH", raising=False)
    else:
        monkeypatch.setenv("HM_WPATH", WPATHS[wpath])
    _say(f"case: products {name} R={R} wpath={wpath}")
    rng = np.random.default_rng(len(leaves) * 131 + R)
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    try:
        dense = ft.to_dense()
        n = ft.n
        x = np.asfortranarray(rng.standard_normal((n, R)))
        dt = D.DeviceTree.upload(t.contents.root, ft.height)
        dx = D.DeviceArray.from_numpy(x)
        dy = D.DeviceArray((n, R))
        if R == 1:
            dt.matvec(dx, dy)
        else:
            dt.matmat(dx, dy)
        assert H.rel_err(dy.numpy(), dense @ x) < TOL
        dt.matmat(dx, dy, trans=1)
        assert H.rel_err(dy.numpy(), dense.T @ x) < TOL
        # X H through the strided transposed form
        xr = np.asfortranarray(rng.standard_normal((R, n)))
        dxr = D.DeviceArray.from_numpy(xr)
        dyr = D.DeviceArray((R, n))
        dt.matmat(dxr, dyr, trans=1, R=R, x_strides=(R, 1), y_strides=(R, 1))
        assert H.rel_err(dyr.numpy(), xr @ dense) < TOL
        dt.free()
    finally:
        T.detach_and_free(lib, t)


@pytest.mark.parametrize("name,leaves,ranks", CASES, ids=[c[0] for c in CASES])
def test_matvec_fused_and_reduced(gpu, monkeypatch, name, leaves, ranks):
    """H x with the chunk partials summed inside the apply kernel (trees
    whose leaves have few partial values) and through the two reduction
    launches (HM_MV_FUSE=0; the path larger trees take): both vs the dense
    product, and against each other."""
    from hmat_lib_amd import device as D
    _say(f"case: matvec fused/reduced {name}")
    rng = np.random.default_rng(len(leaves) * 17 + 3)
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    try:
        dense = ft.to_dense()
        x = rng.standard_normal(ft.n)
        dt = D.DeviceTree.upload(t.contents.root, ft.height)
        dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((ft.n,))
        ys = {}
        for fuse in ("1", "0"):
            monkeypatch.setenv("HM_MV_FUSE", fuse)
            dt.matvec(dx, dy)
            dt.matvec(dx, dy)
            ys[fuse] = dy.numpy()
            assert H.rel_err(ys[fuse], dense @ x) < TOL, fuse
        assert H.rel_err(ys["1"], ys["0"]) < 1e-14
        dt.free()
    finally:
        T.detach_and_free(lib, t)


@pytest.mark.parametrize("variant", range(7))
def test_every_apply_tile_variant(gpu, monkeypatch, variant):
    """Each tile shape of the apply kernel (forced instead of tuned), on a
    uniform tree (bounds-check-free path) and a ragged one, H X and H^T X,
    R = 64 and 100."""
    from hmat_lib_amd import device as D
    monkeypatch.setenv("HM_APPLY_VARIANT", str(variant))
    for name, leaves, ranks in (CASES[1], CASES[2], CASES[4]):
        for R in (64, 100):
            _say(f"case: apply variant {variant} {name} R={R}")
            rng = np.random.default_rng(variant * 7 + R)
            ft = H.random_flat_tree(rng, leaves, ranks)
            lib = capi.hmat()
            keep = T.Keep()
            t = T.fill_tree(lib, ft, keep)
            try:
                dense = ft.to_dense()
                x = np.asfortranarray(rng.standard_normal((ft.n, R)))
                dt = D.DeviceTree.upload(t.contents.root, ft.height)
                dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((ft.n, R))
                dt.matmat(dx, dy)
                assert capi.hip().hm_hip_matmat_variant(dt.h, 0, R) == variant
                assert H.rel_err(dy.numpy(), dense @ x) < TOL
                dt.matmat(dx, dy, trans=1)
                assert H.rel_err(dy.numpy(), dense.T @ x) < TOL
                dt.free()
            finally:
                T.detach_and_free(lib, t)


LEAF_CASES = [
    ("leaf 256 h3 rank 1", [256] * 8, lambda l, j: 1),
    ("leaf 256 h3 ranks to 24 (two low-rank chunks)", [256] * 8, lambda l, j: 24 - 3 * l + j % 2),
    ("leaf 128 h2 some rank 0", [128] * 4, lambda l, j: 0 if (l + j) % 3 == 0 else 5 + j),
    ("leaf 512 h2", [512] * 4, lambda l, j: 7 + l),
]


@pytest.mark.parametrize("variant", [5, 6])
@pytest.mark.parametrize("case", range(len(LEAF_CASES)))
def test_leaf_resident_apply(gpu, monkeypatch, case, variant):
    """The leaf-resident apply kernels (variant 5: one workgroup per leaf,
    D streamed into registers, X through a two-chunk swizzled LDS ring;
    variant 6: X_c resident in LDS, two passes of 128 rows, low-rank K at
    most 32) on uniform trees of leaf 128 / 256 / 512, low-rank K of one and
    several chunks and blocks of rank 0, R = 64 and 192; then operands they
    do not take (R = 100, X with an odd leading dimension, H^T X; for
    variant 6 also leaf 512 and low-rank K above 32), which fall back to
    variant 0."""
    from hmat_lib_amd import device as D
    monkeypatch.setenv("HM_APPLY_VARIANT", str(variant))
    name, leaves, ranks = LEAF_CASES[case]
    rng = np.random.default_rng(40 + case)
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    try:
        dense = ft.to_dense()
        n = ft.n
        dt = D.DeviceTree.upload(t.contents.root, ft.height)
        for R in (64, 192, 100):
            _say(f"case: leaf apply v{variant} {name} R={R}")
            x = np.asfortranarray(rng.standard_normal((n, R)))
            dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((n, R))
            dt.matmat(dx, dy)
            assert capi.hip().hm_hip_matmat_variant(dt.h, 0, R) == variant
            assert H.rel_err(dy.numpy(), dense @ x) < TOL
        _say(f"case: leaf apply {name} odd leading dimension and transpose")
        R = 64
        xp = np.asfortranarray(rng.standard_normal((n + 1, R)))
        dxp = D.DeviceArray.from_numpy(xp)
        dy = D.DeviceArray((n, R))
        dt.matmat(dxp, dy, R=R, x_strides=(1, n + 1))
        assert H.rel_err(dy.numpy(), dense @ xp[:n]) < TOL
        dt.matmat(dxp, dy, trans=1, R=R, x_strides=(1, n + 1))
        assert H.rel_err(dy.numpy(), dense.T @ xp[:n]) < TOL
        dt.free()
    finally:
        T.detach_and_free(lib, t)


def test_dropin_api_vs_oracle(gpu):
    """multiply_* of libhmat_lib vs the oracle's restatement, same tree."""
    _say("case: drop-in API vs oracle")
    rng = np.random.default_rng(7)
    ft = H.random_flat_tree(rng, [50, 40, 33, 60, 45, 41, 39, 52], lambda l, j: 1 + (l * 3 + j) % 4)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    try:
        n = ft.n
        ierr = C.c_int(-1)
        x = rng.standard_normal(n)
        y = np.zeros(n)
        lib.multiply_vector(t, capi.dptr(x), capi.dptr(y), C.byref(ierr))
        assert ierr.value == 0
        assert H.rel_err(y, H.oracle_matvec(t, x)) < TOL
        X = np.asfortranarray(rng.standard_normal((n, 16)))
        for name, mode in (("multiply_hodlr_dense", "hd"),
                           ("multiply_hodlr_transpose_dense", "htd")):
            out = np.zeros((n, 16), order="F")
            getattr(lib, name)(t, capi.dptr(X), 16, n, capi.dptr(out), n, C.byref(ierr))
            assert ierr.value == 0
            assert H.rel_err(out, H.oracle_matmat(t, X, mode)) < TOL
        Xr = np.asfortranarray(rng.standard_normal((9, n)))
        out = np.zeros((9, n), order="F")
        lib.multiply_dense_hodlr(t, capi.dptr(Xr), 9, 9, capi.dptr(out), 9, C.byref(ierr))
        assert ierr.value == 0
        assert H.rel_err(out, H.oracle_matmat(t, Xr, "dh")) < TOL
        # internal-node forms on the root and on a subtree
        out = np.zeros((n, 16), order="F")
        lib.multiply_internal_node_dense(t.contents.root, ft.height, capi.dptr(X), 16, n,
                                         None, None, capi.dptr(out), n)
        assert H.rel_err(out, ft.to_dense() @ X) < TOL
        sub = t.contents.root.contents.children[3].internal
        m3 = sub.contents.m
        Xs = np.asfortranarray(rng.standard_normal((m3, 5)))
        outs = np.zeros((m3, 5), order="F")
        lib.multiply_internal_node_transpose_dense(sub, ft.height - 1, capi.dptr(Xs), 5, m3,
                                                   None, None, capi.dptr(outs), m3)
        off = ft.n - m3
        dsub = ft.to_dense()[off:, off:]
        assert H.rel_err(outs, dsub.T @ Xs) < TOL
        outd = np.zeros((5, m3), order="F")
        Xd = np.asfortranarray(rng.standard_normal((5, m3)))
        lib.multiply_dense_internal_node(sub, ft.height - 1, capi.dptr(Xd), 5, 5,
                                         None, None, capi.dptr(outd), 5)
        assert H.rel_err(outd, Xd @ dsub) < TOL
    finally:
        T.detach_and_free(lib, t)


def test_dropin_errors(gpu):
    lib = capi.hmat()
    ierr = C.c_int(0)
    t = lib.allocate_tree_monolithic(2, C.byref(ierr))
    x = np.zeros(4)
    assert not lib.multiply_vector(None, capi.dptr(x), None, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    assert not lib.multiply_vector(t, capi.dptr(x), capi.dptr(x), C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    lib.free_tree_hodlr(C.byref(t))
