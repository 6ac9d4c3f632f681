"""hm_hip_matmat_noalloc, the allocation-free variant of hm_hip_matmat (its
scratch is a fixed device array of the library), through the C-ABI:

* on trees hm_hip_matmat has tuned, its output is bitwise equal to
  hm_hip_matmat's (same kernels, same apply variant, only the scratch
  differs); on an untuned tree it runs variant 1 and matches exact dense
  algebra to 1e-12;
* two calls in a row, the second with a smaller input (a smaller ragged
  tree, fewer right-hand sides, H^T X and the R = 1 path), each vs
  hm_hip_matmat / hm_hip_matvec;
* free device memory is unchanged across calls on a tree whose own scratch
  was never allocated;
* a product needing more than the fixed scratch, and a tree with sliced
  factors, give INPUT_ERROR before any kernel runs.
"""
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


def _free_device_memory():
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(0), C.c_size_t(0)
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def _noalloc(dt, dx, dy, trans=0, R=None):
    n = dx.shape[0]
    R = R if R is not None else (dx.shape[1] if len(dx.shape) > 1 else 1)
    return capi.hip().hm_hip_matmat_noalloc(dt.h, trans, R, dx.ptr, 1, n, dy.ptr, 1,
                                            dy.shape[0], None)


def _tree(rng, leaves, ranks):
    from hmat_lib_amd import device as D
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()  # owns the arrays behind the host tree's pointers
    t = T.fill_tree(lib, ft, keep)
    dt = D.DeviceTree.upload(t.contents.root, ft.height)
    T.detach_and_free(lib, t)
    del keep
    return ft, dt


def test_noalloc_matches_matmat(gpu):
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(123)
    # first input: uniform leaves of 256 (the leaf-resident apply), R = 64
    ft, dt = _tree(rng, [256] * 16, lambda l, j: 1 + (l + j) % 3)
    ft3, dt3 = _tree(rng, [256] * 16, lambda l, j: 2)
    n, R = ft.n, 64
    dense = ft.to_dense()
    x = np.asfortranarray(rng.standard_normal((n, R)))
    dx, dy, dy2 = D.DeviceArray.from_numpy(x), D.DeviceArray((n, R)), D.DeviceArray((n, R))
    dy3 = D.DeviceArray((n, R))
    _say("case: noalloc, untuned tree, R=64")
    capi.check(_noalloc(dt, dx, dy), "matmat_noalloc")
    D.synchronize()
    assert H.rel_err(dy.numpy(), dense @ x) < TOL
    # the code object is loaded by now: calls on a tree whose own scratch was
    # never allocated leave the free device memory unchanged
    free0 = _free_device_memory()
    for r in (64, 5, 1):
        capi.check(_noalloc(dt3, dx, dy3, R=r), "matmat_noalloc")
    D.synchronize()
    assert _free_device_memory() == free0, "the allocation-free product allocated"
    assert H.rel_err(dy3.numpy()[:, :1], ft3.to_dense() @ x[:, :1]) < TOL
    _say("case: noalloc vs matmat after tuning, R=64")
    dt.matmat(dx, dy2)
    capi.check(_noalloc(dt, dx, dy), "matmat_noalloc")
    D.synchronize()
    np.testing.assert_array_equal(dy.numpy(), dy2.numpy())

    # second input, smaller: a ragged tree, fewer right-hand sides, H^T X
    ft2, dt2 = _tree(rng, [90, 33, 1, 70, 41, 0, 17, 60], lambda l, j: 1 + (5 * l + j) % 7)
    n2 = ft2.n
    dense2 = ft2.to_dense()
    for trans in (0, 1):
        for R2 in (5, 1):
            _say(f"case: noalloc, smaller ragged tree, R={R2}, trans={trans}")
            x2 = np.asfortranarray(rng.standard_normal((n2, R2)))
            ex, ey, ey2 = (D.DeviceArray.from_numpy(x2), D.DeviceArray((n2, R2)),
                           D.DeviceArray((n2, R2)))
            if R2 == 1 and trans == 0:
                dt2.matvec(ex, ey2)
            else:
                dt2.matmat(ex, ey2, trans=trans)
            capi.check(_noalloc(dt2, ex, ey, trans=trans), "matmat_noalloc")
            D.synchronize()
            np.testing.assert_array_equal(ey.numpy(), ey2.numpy())
            ref = (dense2.T if trans else dense2) @ x2
            assert H.rel_err(ey.numpy(), ref) < TOL

    _say("case: noalloc, scratch limit and sliced factors refused")
    big = 20000  # partials + W of tree 1 at this R exceed the fixed scratch
    rc = capi.hip().hm_hip_matmat_noalloc(dt.h, 0, big, dx.ptr, 1, n, dy.ptr, 1, n, None)
    assert rc == capi.INPUT_ERROR
    assert "scratch" in capi.last_error()
    dt.set_precision(2)
    assert _noalloc(dt, dx, dy) == capi.INPUT_ERROR
    dt.set_precision(0)
    capi.check(_noalloc(dt, dx, dy), "matmat_noalloc")
    D.synchronize()
    np.testing.assert_array_equal(dy.numpy(), dy2.numpy())
    dt.free()
    dt2.free()
    dt3.free()
