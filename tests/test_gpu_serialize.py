"""Binary HODLR files on the device side: a device tree saved by
hm_hip_tree_save is byte-identical to the same tree saved from the host,
and a file loads as a device tree whose products and factors equal the
original's (bit for bit: the pools are copied, not recomputed)."""
import ctypes as C
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from hmat_lib_amd import workloads as WL
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_device_and_host_files_agree(gpu, tmp_path):
    from hmat_lib_amd import device as D
    _say("case: serialize random tree")
    rng = np.random.default_rng(8)
    leaves = [9, 0, 14, 6, 11, 3, 8, 7]
    ft = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(0, 5)))
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    host_file, dev_file = tmp_path / "host.hodlr", tmp_path / "dev.hodlr"
    assert lib.hmat_amd_save_tree(t, str(host_file).encode()) == capi.SUCCESS
    dt = D.DeviceTree.upload(t.contents.root, 3)
    dt.save(str(dev_file))
    assert host_file.read_bytes() == dev_file.read_bytes()
    # host file -> device tree: same products as the original
    d2 = D.DeviceTree.load(str(host_file))
    x = np.asfortranarray(rng.standard_normal((ft.n, 5)))
    dx, y1, y2 = D.DeviceArray.from_numpy(x), D.DeviceArray((ft.n, 5)), D.DeviceArray((ft.n, 5))
    dt.matmat(dx, y1)
    d2.matmat(dx, y2)
    np.testing.assert_array_equal(y1.numpy(), y2.numpy())
    assert H.rel_err(y2.numpy(), ft.to_dense() @ x) < 1e-12
    for d in (dt, d2):
        d.free()
    T.detach_and_free(lib, t)


def test_compressed_tree_round_trip(gpu, tmp_path):
    """A device-compressed exp1d tree (N = 2048): device file -> host tree
    equals the downloaded original; device file -> device tree reproduces
    the products exactly."""
    from hmat_lib_amd import device as D
    _say("case: serialize compressed exp1d n=2048")
    n, h = 2048, 4
    pts = np.sort(WL.uniform(31, n))[:, None]
    da = D.fill_kernel_matrix(0, pts, 0.05, 0.0)
    dt = D.DeviceTree.compress(da, n, h, 1e-10)
    path = tmp_path / "c.hodlr"
    dt.save(str(path))
    lib = capi.hmat()
    ierr = C.c_int(0)
    ref = lib.allocate_tree_monolithic(h, C.byref(ierr))
    dt.download(lib, ref)
    loaded = lib.hmat_amd_load_tree(str(path).encode(), C.byref(ierr))
    assert ierr.value == capi.SUCCESS
    fa, fb = T.flatten(ref), T.flatten(loaded)
    for d1, d2 in zip(fa.diag, fb.diag):
        np.testing.assert_array_equal(d1, d2)
    for k in fa.blocks:
        np.testing.assert_array_equal(fa.blocks[k][0], fb.blocks[k][0])
        np.testing.assert_array_equal(fa.blocks[k][1], fb.blocks[k][1])
    d2 = D.DeviceTree.load(str(path))
    x = np.asfortranarray(np.random.default_rng(1).standard_normal((n, 3)))
    dx, y1, y2 = D.DeviceArray.from_numpy(x), D.DeviceArray((n, 3)), D.DeviceArray((n, 3))
    dt.matmat(dx, y1)
    d2.matmat(dx, y2)
    np.testing.assert_array_equal(y1.numpy(), y2.numpy())
    lib.free_tree_hodlr(C.byref(ref))
    lib.free_tree_hodlr(C.byref(loaded))
    for d in (dt, d2):
        d.free()
    da.free()


def test_device_load_rejects_truncated_file(gpu, tmp_path):
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(9)
    ft = H.random_flat_tree(rng, [5, 6, 4, 7], lambda l, j: 2)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    p = tmp_path / "x.hodlr"
    assert lib.hmat_amd_save_tree(t, str(p).encode()) == capi.SUCCESS
    T.detach_and_free(lib, t)
    p.write_bytes(p.read_bytes()[:-16])
    out = C.c_void_p()
    rc = capi.hip().hm_hip_tree_load(str(p).encode(), C.byref(out), None)
    assert rc == capi.INPUT_ERROR and not out.value
