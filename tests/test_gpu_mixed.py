"""Mixed-precision products: off-diagonal factors split into 1..4 BF16
slices (exact FP32 accumulation) or 1..4 INT8 slices (exact int32
accumulation) on MFMA, or one FP32 copy (FP32 accumulation), diagonal leaves
FP64, vs the FP64 dense product of the same tree — for H X, H^T X, X H
(strided operands), H x (R = 1), and for batches of trees with different
formats and slice counts beside FP64 trees.

Tolerance per slice count (relative, Frobenius): each BF16 or INT8 slice
carries 8 bits of the chunk maximum, so 1 slice 3e-2, 2 slices 5e-4, 3
slices 2e-6, 4 slices 1e-8 (the workload target); FP32 1e-6 (operands and
sums rounded to 24 bits).  Trees cover ragged and empty leaves, leaves
longer than one chunk (128 rows for BF16, 256 for INT8, 32 for FP32) and
ranks spanning several rank chunks."""
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from tests import helpers as H

pytestmark = pytest.mark.gpu
TOLS = {1: 3e-2, 2: 5e-4, 3: 2e-6, 4: 1e-8}
TOL_FP32 = 1e-6
# (format, slice counts) of the products tested
FORMATS = {"bf16": (1, 2, 3, 4), "int8": (1, 2, 3, 4), "fp32": (1,)}


def _tol(fmt, S):
    return TOL_FP32 if fmt == "fp32" else TOLS[S]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ("uniform h3", [128] * 8, lambda l, j: 5 + 3 * l),
    ("ragged zeros h3", [200, 0, 37, 129, 64, 300, 1, 90], lambda l, j: 1 + (5 * l + j) % 23),
    ("rank 150 h2", [260, 250, 240, 270], lambda l, j: 150 if l == 1 else 70),
    ("rank 300 h1", [700, 650], lambda l, j: 300),
]


def _upload(rng, leaves, ranks):
    from hmat_lib_amd import device as D
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    dt = D.DeviceTree.upload(t.contents.root, ft.height)
    T.detach_and_free(lib, t)
    del keep
    return ft, dt


@pytest.mark.parametrize("fmt", list(FORMATS))
@pytest.mark.parametrize("name,leaves,ranks", CASES, ids=[c[0] for c in CASES])
def test_mixed_products_vs_dense(gpu, name, leaves, ranks, fmt):
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(len(leaves) + sum(leaves))
    ft, dt = _upload(rng, leaves, ranks)
    dense = ft.to_dense()
    n = ft.n
    for R in (3, 64, 100):
        x = np.asfortranarray(rng.standard_normal((n, R)))
        dx = D.DeviceArray.from_numpy(x)
        dy = D.DeviceArray((n, R))
        xr = np.asfortranarray(rng.standard_normal((R, n)))
        dxr, dyr = D.DeviceArray.from_numpy(xr), D.DeviceArray((R, n))
        for S in FORMATS[fmt]:
            _say(f"case: mixed {name} {fmt} R={R} slices={S}")
            dt.set_precision(S, fmt)
            tol = _tol(fmt, S)
            dt.matmat(dx, dy)
            dt.matmat(dx, dy)  # cached program, same result
            err = H.rel_err(dy.numpy(), dense @ x)
            assert err < tol, ("H X", S, err)
            dt.matmat(dx, dy, trans=1)
            err = H.rel_err(dy.numpy(), dense.T @ x)
            assert err < tol, ("H^T X", S, err)
            # X H through the strided transposed form
            dt.matmat(dxr, dyr, trans=1, R=R, x_strides=(R, 1), y_strides=(R, 1))
            err = H.rel_err(dyr.numpy(), xr @ dense)
            assert err < tol, ("X H", S, err)
        dt.set_precision(0)
        dt.matmat(dx, dy)
        assert H.rel_err(dy.numpy(), dense @ x) < 1e-12
    # matvec (R = 1) on sliced factors
    xv = rng.standard_normal(n)
    dxv, dyv = D.DeviceArray.from_numpy(xv), D.DeviceArray((n,))
    for S in FORMATS[fmt][1::2] or FORMATS[fmt]:
        _say(f"case: mixed {name} {fmt} matvec slices={S}")
        dt.set_precision(S, fmt)
        dt.matvec(dxv, dyv)
        assert H.rel_err(dyv.numpy(), dense @ xv) < _tol(fmt, S)
    dt.set_precision(0)
    dt.free()


def test_set_format_rejects_bad_input(gpu):
    """Unknown formats and slice counts outside a format's range are refused
    and leave the tree's products as they were."""
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(5)
    ft, dt = _upload(rng, [64] * 4, lambda l, j: 3)
    x = np.asfortranarray(rng.standard_normal((ft.n, 8)))
    dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((ft.n, 8))
    lib = capi.hip()
    for code, S in ((4, 1), (-1, 1), (3, 2), (2, 5), (1, 0), (2, 0)):
        assert lib.hm_hip_tree_set_format(dt.h, code, S, None) == capi.INPUT_ERROR, (code, S)
    dt.matmat(dx, dy)
    assert H.rel_err(dy.numpy(), ft.to_dense() @ x) < 1e-12
    dt.free()


@pytest.mark.parametrize("trans", [0, 1, 2])
def test_mixed_batch(gpu, trans):
    """A batch of uneven trees with BF16, INT8 and FP32 factor copies of
    several slice counts and FP64 trees: the FP64 diagonal chains plus one
    launch per sliced stage for all trees of one format and slice count; vs dense per tree at each tree's tolerance, tuned and from HIP
    graph replays; a tree whose precision changes afterwards is refused."""
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(60 + trans)
    specs = [([128] * 8, 4, "bf16"), ([90, 33, 1, 70, 41, 0, 17, 60], 2, "bf16"),
             ([200, 150, 0, 64], 0, "fp64"), ([256] * 4, 4, "int8"), ([64] * 16, 0, "fp64"),
             ([300, 129, 77, 256], 4, "bf16"), ([100, 300, 0, 60], 1, "fp32"),
             ([128] * 4, 2, "int8"), ([520, 7, 33, 260], 4, "int8")]
    trees, dense, slices = [], [], []
    for leaves, S, fmt in specs:
        ft, dt = _upload(rng, leaves, lambda l, j: 1 + (7 * l + 3 * j) % 30)
        if S:
            dt.set_precision(S, fmt)
        trees.append(dt)
        dense.append(ft.to_dense())
        slices.append(_tol(fmt, S) if S else 1e-12)
    xs, ys, xh = [], [], []
    for i, d in enumerate(dense):
        n = d.shape[0]
        R = [1, 5, 64, 100, 17, 70, 33, 64, 9][i]
        shape = (R, n) if trans == 2 else (n, R)
        x = np.asfortranarray(rng.standard_normal(shape))
        xh.append(x)
        xs.append(D.DeviceArray.from_numpy(x))
        ys.append(D.DeviceArray(shape))
    b = D.Batch(trees, xs, ys, trans=trans)

    def check(tag):
        for i, d in enumerate(dense):
            ref = (xh[i] @ d if trans == 2 else (d.T if trans else d) @ xh[i])
            err = H.rel_err(ys[i].numpy(), ref)
            assert err < slices[i], (tag, i, slices[i], err)

    _say(f"case: mixed batch trans={trans}")
    b.run()
    D.synchronize()
    check("eager")
    b.tune()
    for rep in range(3):
        b.run(graph=True)
        D.synchronize()
        check(f"graph {rep}")
    trees[0].set_precision(3)
    with pytest.raises(RuntimeError, match="precision"):
        b.run()
    b.free()
    for t in trees:
        t.free()
