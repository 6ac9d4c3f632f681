"""Walking and extracting HODLR trees of the hmat_lib ABI from Python.

Works on any TreeHODLR* (product or checker library).  Block order used
throughout: level l = 1..h, index j = 0..2^l-1, j even = top-right child [1]
of parent j//2, j odd = bottom-left child [2]; block (l, j) covers the rows
of node (l, j) and the columns of node (l, j ^ 1) (this is also the device
tree's order).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi


def levels(node_ptr, height):
    """Internal nodes per depth d = 0..height-1, left to right."""
    out = [[node_ptr]]
    for _ in range(height - 1):
        nxt = []
        for p in out[-1]:
            nxt.append(p.contents.children[0].internal)
            nxt.append(p.contents.children[3].internal)
        out.append(nxt)
    return out


def _arr(ptr, count):
    if count <= 0:
        return np.zeros(0)
    if not ptr:
        raise ValueError("NULL data pointer in tree")
    return np.ctypeslib.as_array(ptr, shape=(count,)).copy()


@dataclass
class FlatTree:
    """Host copy of a HODLR tree as numpy arrays."""
    height: int
    leaf_sizes: list
    diag: list                     # m_i x m_i arrays
    blocks: dict = field(default_factory=dict)  # (l, j) -> (U m x s, V n x s)

    @property
    def n(self):
        return int(sum(self.leaf_sizes))

    def leaf_off(self):
        return np.concatenate([[0], np.cumsum(self.leaf_sizes)]).astype(np.int64)

    def node_range(self, level, q):
        off = self.leaf_off()
        span = 1 << (self.height - level)
        return int(off[q * span]), int(off[(q + 1) * span])

    def block_ranges(self, level, j):
        r = self.node_range(level, j)
        c = self.node_range(level, j ^ 1)
        return r, c

    def ranks(self):
        return {k: v[0].shape[1] for k, v in self.blocks.items()}

    def to_dense(self):
        n = self.n
        a = np.zeros((n, n))
        off = self.leaf_off()
        for i, d in enumerate(self.diag):
            a[off[i]:off[i + 1], off[i]:off[i + 1]] = d
        for (l, j), (u, v) in self.blocks.items():
            (r0, r1), (c0, c1) = self.block_ranges(l, j)
            a[r0:r1, c0:c1] = u @ v.T
        return a


def flatten(tree_ptr) -> FlatTree:
    """Copy a TreeHODLR (pointer) into a FlatTree."""
    t = tree_ptr.contents
    h = t.height
    lev = levels(t.root, h)
    leaf_sizes, diag = [], []
    for p in lev[-1]:
        for c in (0, 3):
            d = p.contents.children[c].leaf.contents.data.diagonal
            leaf_sizes.append(d.m)
            diag.append(_arr(d.data, d.m * d.m).reshape(d.m, d.m, order="F"))
    ft = FlatTree(h, leaf_sizes, diag)
    for l in range(1, h + 1):
        for j in range(1 << l):
            p = lev[l - 1][j >> 1]
            od = p.contents.children[2 if (j & 1) else 1].leaf.contents.data.off_diagonal
            u = _arr(od.u, od.m * od.s).reshape(od.m, od.s, order="F")
            v = _arr(od.v, od.n * od.s).reshape(od.n, od.s, order="F")
            ft.blocks[(l, j)] = (u, v)
    return ft


class Keep:
    """Holds numpy buffers alive while a tree points into them."""

    def __init__(self):
        self.bufs = []

    def add(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        self.bufs.append(a)
        return a


def fill_tree(lib, flat: FlatTree, keep: Keep):
    """Allocate a tree with `lib` and point its fields at copies of `flat`'s
    arrays (kept alive by `keep`; do not free_tree_data such a tree).
    Returns the TreeHODLR pointer."""
    ierr = C.c_int(0)
    t = lib.allocate_tree(flat.height, C.byref(ierr))
    if ierr.value != capi.SUCCESS:
        raise RuntimeError("allocate_tree failed")
    h = flat.height
    lev = levels(t.contents.root, h)
    off = flat.leaf_off()
    for d in range(h):
        span = 1 << (h - d)
        for q, p in enumerate(lev[d]):
            p.contents.m = int(off[(q + 1) * span] - off[q * span])
    i = 0
    for p in lev[-1]:
        for c in (0, 3):
            dg = p.contents.children[c].leaf.contents.data.diagonal
            a = keep.add(flat.diag[i])
            dg.m = a.shape[0]
            dg.data = a.ctypes.data_as(capi.PDouble)
            i += 1
    for (l, j), (u, v) in flat.blocks.items():
        p = lev[l - 1][j >> 1]
        od = p.contents.children[2 if (j & 1) else 1].leaf.contents.data.off_diagonal
        uu, vv = keep.add(u), keep.add(v)
        od.m, od.n, od.s = uu.shape[0], vv.shape[0], uu.shape[1]
        od.u = uu.ctypes.data_as(capi.PDouble)
        od.v = vv.ctypes.data_as(capi.PDouble)
    return t


def detach_and_free(lib, t):
    """Free a tree made by fill_tree without freeing numpy-owned arrays."""
    h = t.contents.height
    lev = levels(t.contents.root, h)
    for d in range(h):
        for p in lev[d]:
            for c in (1, 2):
                od = p.contents.children[c].leaf.contents.data.off_diagonal
                od.u = None
                od.v = None
    for p in lev[-1]:
        for c in (0, 3):
            p.contents.children[c].leaf.contents.data.diagonal.data = None
    lib.free_tree_hodlr(C.byref(t))
