"""Batched products over many uneven trees (sizes, heights, ranks and RHS
counts all differ, zero-size leaves included) and batched compression, vs
exact dense algebra and vs the one-tree path.

Tolerances: products 1e-12 relative (FP64, only the summation order
differs); compressed products vs the FP64 dense product 1e-8 relative."""
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from hmat_lib_amd import workloads as WL
from tests import helpers as H

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _forest(seed, count):
    """(flat tree, device tree, R) for `count` random uneven trees."""
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(seed)
    lib = capi.hmat()
    out = []
    for i in range(count):
        h = int(rng.integers(1, 5))
        leaves = [int(v) for v in rng.integers(0 if i % 4 == 0 else 1, 90, size=1 << h)]
        if sum(leaves) == 0:
            leaves[0] = 5
        smax = int(rng.integers(1, 25))
        ft = H.random_flat_tree(rng, leaves, lambda l, j, smax=smax: 1 + (7 * l + 3 * j) % smax)
        keep = T.Keep()
        t = T.fill_tree(lib, ft, keep)
        dt = D.DeviceTree.upload(t.contents.root, ft.height)
        T.detach_and_free(lib, t)
        R = int([1, 2, 5, 16, 64, 100][i % 6])
        out.append((ft, dt, R))
    return out, rng


ROUTES = ["tree", "chains"]


def _route(monkeypatch, route):
    """"tree": trees with the chunk W path run the batched tree-product
    kernels (partials, reductions, apply tiles), the rest the chained GEMMs;
    "chains": every tree the chained GEMMs."""
    if route == "chains":
        monkeypatch.setenv("HM_BATCH_CHAINS", "1")
    else:
        monkeypatch.delenv("HM_BATCH_CHAINS", raising=False)


@pytest.mark.parametrize("route", ROUTES)
@pytest.mark.parametrize("trans", [0, 1])
def test_batch_products_vs_dense(gpu, trans, route, monkeypatch):
    from hmat_lib_amd import device as D
    _route(monkeypatch, route)
    _say(f"case: batch of 14 uneven trees trans={trans} route={route}")
    forest, rng = _forest(7 + trans, 14)
    xs, ys, dense, xh = [], [], [], []
    for ft, dt, R in forest:
        x = np.asfortranarray(rng.standard_normal((ft.n, R)))
        xh.append(x)
        xs.append(D.DeviceArray.from_numpy(x))
        ys.append(D.DeviceArray((ft.n, R)))
        d = ft.to_dense()
        dense.append(d.T if trans else d)
    b = D.Batch([f[1] for f in forest], xs, ys, trans=trans)

    def check():
        for i in range(len(forest)):
            assert H.rel_err(ys[i].numpy(), dense[i] @ xh[i]) < TOL, i

    b.run()
    D.synchronize()
    check()
    b.tune()
    v = b.variants()
    assert len(v) >= 2 and all(0 <= k < 6 for k in v)
    b.run()
    D.synchronize()
    check()
    eager = [y.numpy() for y in ys]
    # graph replays, with eager work and synchronisation in between; new X
    # contents between replays must be picked up
    scratch = D.DeviceArray(ys[3].shape)
    for rep in range(3):
        b.run(graph=True)
        D.synchronize()
        for i in range(len(forest)):
            np.testing.assert_array_equal(ys[i].numpy(), eager[i])
        forest[3][1].matmat(xs[3], scratch, trans=trans)
        D.synchronize()
        assert H.rel_err(scratch.numpy(), eager[3]) < TOL
    xh[2] = np.asfortranarray(rng.standard_normal(xh[2].shape))
    xs[2].upload(xh[2])
    b.run(graph=True)
    D.synchronize()
    check()
    b.free()


@pytest.mark.parametrize("variant", range(6))
def test_batch_apply_variants(gpu, variant, monkeypatch):
    """Every apply tile shape of the batched tree products (64 x 64,
    128 x 64, 128 x 32, 64 x 32, 64 x 16 on four waves; 32 x 32 on one),
    fixed with HM_BATCH_APPLY_VARIANT,
    on ragged leaves (zero-size ones included) and R = 1 .. 100, vs dense
    algebra; H X and H^T X."""
    from hmat_lib_amd import device as D
    _route(monkeypatch, "tree")
    monkeypatch.setenv("HM_BATCH_APPLY_VARIANT", str(variant))
    for trans in (0, 1):
        _say(f"case: batch apply variant {variant} trans={trans}")
        forest, rng = _forest(31 + trans, 12)
        xs, ys, xh, dense = [], [], [], []
        for ft, dt, R in forest:
            x = np.asfortranarray(rng.standard_normal((ft.n, R)))
            xh.append(x)
            xs.append(D.DeviceArray.from_numpy(x))
            ys.append(D.DeviceArray((ft.n, R)))
            d = ft.to_dense()
            dense.append(d.T if trans else d)
        b = D.Batch([f[1] for f in forest], xs, ys, trans=trans)
        b.run()
        D.synchronize()
        for i in range(len(forest)):
            assert H.rel_err(ys[i].numpy(), dense[i] @ xh[i]) < TOL, (trans, i)
        b.free()


@pytest.mark.parametrize("route", ROUTES)
def test_batch_dense_times_hodlr(gpu, route, monkeypatch):
    """Batched X_i H_i (X_i, Y_i: R_i x n_i, the batched form of
    multiply_dense_hodlr): taken as Y^T = H^T X^T with transposed operand
    access and transposed stores; untuned, tuned and replayed from a graph."""
    from hmat_lib_amd import device as D
    _route(monkeypatch, route)
    _say(f"case: batch of 14 uneven trees, X H, route={route}")
    forest, rng = _forest(23, 14)
    xs, ys, xh, dense = [], [], [], []
    for ft, dt, R in forest:
        x = np.asfortranarray(rng.standard_normal((R, ft.n)))
        xh.append(x)
        xs.append(D.DeviceArray.from_numpy(x))
        ys.append(D.DeviceArray((R, ft.n)))
        dense.append(ft.to_dense())
    b = D.Batch([f[1] for f in forest], xs, ys, trans=2)

    def check():
        for i in range(len(forest)):
            assert H.rel_err(ys[i].numpy(), xh[i] @ dense[i]) < TOL, i

    b.run()
    D.synchronize()
    check()
    b.tune()
    b.run(graph=True)
    D.synchronize()
    check()
    b.free()


def test_batch_compress_exp1d(gpu):
    """Six sorted 1-D exponential-kernel matrices of uneven size in one batched
    compression: every block is exactly rank 1 (closed form), ranks equal the
    one-matrix path's, products match the FP64 dense product."""
    from hmat_lib_amd import device as D
    sizes = [256, 640, 1024, 1536, 2048, 3200]
    _say(f"case: batch compress {sizes}")
    mats, dens, hs, pts_l = [], [], [], []
    for i, n in enumerate(sizes):
        pts = WL.points_1d_sorted(900 + i, n)
        ell = 0.02 + 0.08 * i
        da = D.fill_kernel_matrix(0, pts, ell, 0.0)
        mats.append(da)
        dens.append(da.numpy())
        hs.append(max(1, int(round(np.log2(n / 128)))))
    trees, st = D.compress_batch(mats, hs, 1e-8)
    assert st.sketch_rounds >= 1
    for i, n in enumerate(sizes):
        r = trees[i].ranks()
        assert (r == 1).all(), (i, r)
        single = D.DeviceTree.compress(mats[i], n, hs[i], 1e-8)
        np.testing.assert_array_equal(single.ranks(), r)
        R = 3
        x = np.asfortranarray(np.random.default_rng(i).standard_normal((n, R)))
        dx = D.DeviceArray.from_numpy(x)
        dy = D.DeviceArray((n, R))
        trees[i].matmat(dx, dy)
        assert H.rel_err(dy.numpy(), dens[i] @ x) < 1e-8
        single.free()
    # the batched product over the batch-compressed trees
    xs = [D.DeviceArray.from_numpy(np.asfortranarray(
        np.random.default_rng(50 + i).standard_normal((n, 1 + i)))) for i, n in enumerate(sizes)]
    ys = [D.DeviceArray((n, 1 + i)) for i, n in enumerate(sizes)]
    b = D.Batch(trees, xs, ys)
    b.run(graph=True)
    D.synchronize()
    for i in range(len(sizes)):
        assert H.rel_err(ys[i].numpy(), dens[i] @ xs[i].numpy()) < 1e-8
    b.free()
    for t in trees:
        t.free()


def test_batch_compress_shared_pools(gpu):
    """The trees of one batched compression share their device pools: free
    every other tree, the others' products still match the FP64 dense
    product, and the pools go with the last tree."""
    from hmat_lib_amd import device as D
    sizes = [384, 512, 768, 1024, 1280, 640]
    _say(f"case: shared pools of a batch compression {sizes}")
    mats, dens, hs = [], [], []
    for i, n in enumerate(sizes):
        pts = WL.points_1d_sorted(1700 + i, n)
        da = D.fill_kernel_matrix(0, pts, 0.05 + 0.03 * i, 0.0)
        mats.append(da)
        dens.append(da.numpy())
        hs.append(max(1, int(round(np.log2(n / 128)))))
    trees, _ = D.compress_batch(mats, hs, 1e-8)
    for a in mats:
        a.free()
    for i in (0, 2, 4):
        trees[i].free()
    D.synchronize()
    for i in (1, 3, 5):
        n = sizes[i]
        x = np.asfortranarray(np.random.default_rng(i).standard_normal((n, 4)))
        dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((n, 4))
        trees[i].matmat(dx, dy)
        D.synchronize()
        assert H.rel_err(dy.numpy(), dens[i] @ x) < 1e-8, i
    for i in (1, 3, 5):
        trees[i].free()
    D.synchronize()
