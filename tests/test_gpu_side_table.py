"""The drop-in's cache of device trees (side table): products follow the host
tree after its arrays are replaced (free + malloc + refill, usually at the
same address) without any call, and after an in-place edit followed by
hmat_amd_invalidate.  Tolerance 1e-12 relative vs the dense product of the
edited tree."""
import ctypes as C
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _product(lib, t, x):
    n, r = x.shape
    out = np.zeros((n, r), order="F")
    ierr = C.c_int(-1)
    lib.multiply_hodlr_dense(t, capi.dptr(x), r, n, capi.dptr(out), n, C.byref(ierr))
    assert ierr.value == 0, capi.last_error()
    return out


def _off_diag(t, depth_index=0, child=1):
    node = T.levels(t.contents.root, t.contents.height)[0][depth_index]
    return node.contents.children[child].leaf.contents.data.off_diagonal


def test_side_table_tracks_host_edits(gpu):
    _say("case: side table, replaced and edited arrays")
    libc = C.CDLL("libc.so.6")
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]
    lib = capi.hmat()
    rng = np.random.default_rng(31)
    ft = H.random_flat_tree(rng, [40, 37, 50, 33], lambda l, j: 2 + l)
    n = ft.n
    a = np.asfortranarray(ft.to_dense())
    ierr = C.c_int(-1)
    t = lib.allocate_tree_monolithic(2, C.byref(ierr))
    work = a.copy(order="F")
    lib.dense_to_tree_hodlr(t, n, (C.c_int * 4)(40, 37, 50, 33), capi.dptr(work), 1e-14,
                            C.byref(ierr))
    assert ierr.value == 0
    x = np.asfortranarray(rng.standard_normal((n, 5)))
    assert H.rel_err(_product(lib, t, x), a @ x) < 1e-12
    cached = lib.hmat_amd_cache_size()

    # 1. replace U of the root's top-right block by a new array of the same
    #    size with new contents (no invalidate call)
    o = _off_diag(t)
    m, s = o.m, o.s
    old_addr = C.cast(o.u, C.c_void_p).value
    libc.free(old_addr)
    new_addr = libc.malloc(8 * m * s)
    _say(f"  U reallocated at the {'same' if new_addr == old_addr else 'another'} address")
    new_u = rng.standard_normal((m, s))
    C.memmove(new_addr, np.asfortranarray(new_u).ctypes.data, 8 * m * s)
    o.u = C.cast(new_addr, capi.PDouble)
    dense = T.flatten(t).to_dense()
    assert H.rel_err(_product(lib, t, x), dense @ x) < 1e-12
    assert lib.hmat_amd_cache_size() == cached

    # 2. in-place edit of one diagonal entry that no sample covers, then
    #    hmat_amd_invalidate
    leaf = t.contents.innermost_leaves[1].contents.data.diagonal
    mm = leaf.m
    k = mm * (mm // 2) + 3  # an interior entry of column mm // 2
    leaf.data[k] += 10.0
    lib.hmat_amd_invalidate(t)
    assert lib.hmat_amd_cache_size() == cached - 1
    dense = T.flatten(t).to_dense()
    assert H.rel_err(_product(lib, t, x), dense @ x) < 1e-12

    # 3. in-place edit of a sampled entry (the first of a V) is seen without
    #    a call
    o2 = _off_diag(t, 0, 2)
    o2.v[0] = o2.v[0] * -3.0 + 1.0
    dense = T.flatten(t).to_dense()
    assert H.rel_err(_product(lib, t, x), dense @ x) < 1e-12
    lib.free_tree_hodlr(C.byref(t))
