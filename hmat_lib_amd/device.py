"""Device-resident path: thin Python wrappers over the hm_hip C-ABI.

No torch: device memory comes from hm_hip_malloc, copies from hm_hip_memcpy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi


def init(device: int = 0):
    capi.check(capi.hip().hm_hip_init(device), "hm_hip_init")


def device_count() -> int:
    n = C.c_int(0)
    rc = capi.hip().hm_hip_device_count(C.byref(n))
    return n.value if rc == capi.SUCCESS else 0


def synchronize():
    capi.check(capi.hip().hm_hip_synchronize(None), "synchronize")


class DeviceArray:
    """A float64 device buffer of a given shape (column-major semantics)."""

    def __init__(self, shape, zero=False):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.size = int(np.prod(self.shape)) if self.shape else 1
        self.ptr = C.c_void_p()
        capi.check(capi.hip().hm_hip_malloc(C.byref(self.ptr), max(8, self.size * 8)),
                   "hm_hip_malloc")
        if zero:
            self.upload(np.zeros(self.shape, order="F"))

    @classmethod
    def from_numpy(cls, a):
        a = np.asfortranarray(a, dtype=np.float64)
        d = cls(a.shape)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.asfortranarray(a, dtype=np.float64)
        assert a.size == self.size
        capi.check(capi.hip().hm_hip_memcpy_h2d(self.ptr, a.ctypes.data, a.size * 8, None),
                   "h2d")

    def numpy(self):
        out = np.empty(self.shape, order="F")
        capi.check(capi.hip().hm_hip_memcpy_d2h(out.ctypes.data, self.ptr, self.size * 8, None),
                   "d2h")
        return out

    def free(self):
        if self.ptr:
            capi.hip().hm_hip_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def fill_kernel_matrix(kind: int, pts: np.ndarray, ell: float, shift: float = 0.0,
                       out: DeviceArray | None = None) -> DeviceArray:
    """A(i, j) = k(|p_i - p_j|) + shift * delta_ij on the device.
    kind 0 exp(-r/ell), 1 1/(r+ell), 2 Matern-3/2."""
    pts = np.ascontiguousarray(pts, dtype=np.float64)
    n = pts.shape[0]
    dim = 1 if pts.ndim == 1 else pts.shape[1]
    dp = DeviceArray.from_numpy(pts.reshape(-1))
    a = out if out is not None else DeviceArray((n, n))
    capi.check(capi.hip().hm_hip_fill_kernel_matrix(kind, n, dim, dp.ptr, ell, shift,
                                                    a.ptr, n, None), "fill_kernel_matrix")
    synchronize()
    dp.free()
    return a


def default_compress_opts() -> capi.CompressOpts:
    o = capi.CompressOpts()
    capi.hip().hm_hip_compress_opts_default(C.byref(o))
    return o


class DeviceTree:
    """An hm_dtree* owned by Python (destroyed on free/__del__)."""

    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        self.stats = None

    @classmethod
    def upload(cls, node_ptr, height):
        out = C.c_void_p()
        capi.check(capi.hip().hm_hip_tree_upload(node_ptr, height, C.byref(out), None),
                   "tree upload")
        return cls(out)

    @classmethod
    def compress(cls, d_a: DeviceArray, n: int, height: int, thr: float, ms=None,
                 opts: capi.CompressOpts | None = None, lda: int | None = None):
        out = C.c_void_p()
        st = capi.CompressStats()
        ms_arr = None
        if ms is not None:
            ms_arr = (C.c_int * len(ms))(*[int(v) for v in ms])
        rc = capi.hip().hm_hip_compress_dense(
            d_a.ptr, lda or n, n, height, ms_arr, thr,
            C.byref(opts) if opts is not None else None, C.byref(out), C.byref(st), None)
        capi.check(rc, "compress_dense")
        t = cls(out)
        t.stats = st
        return t

    def info(self) -> capi.TreeInfo:
        inf = capi.TreeInfo()
        capi.check(capi.hip().hm_hip_tree_info(self.h, C.byref(inf)), "tree_info")
        return inf

    def ranks(self) -> np.ndarray:
        inf = self.info()
        r = (C.c_int * inf.nblock)()
        capi.check(capi.hip().hm_hip_tree_ranks(self.h, r), "tree_ranks")
        return np.array(r[:], dtype=np.int64)

    def matvec(self, dx: DeviceArray, dy: DeviceArray):
        capi.check(capi.hip().hm_hip_matvec(self.h, dx.ptr, dy.ptr, None), "matvec")

    def matmat(self, dX: DeviceArray, dY: DeviceArray, trans=0, R=None,
               x_strides=None, y_strides=None):
        n = dX.shape[0]
        R = R if R is not None else (dX.shape[1] if len(dX.shape) > 1 else 1)
        xs = x_strides or (1, n)
        ys = y_strides or (1, dY.shape[0])
        capi.check(capi.hip().hm_hip_matmat(self.h, trans, R, dX.ptr, xs[0], xs[1],
                                            dY.ptr, ys[0], ys[1], None), "matmat")

    def save(self, path: str):
        """Write the tree in the binary HODLR layout (hm_hip_tree_save)."""
        capi.check(capi.hip().hm_hip_tree_save(self.h, path.encode(), None), "tree_save")

    @classmethod
    def load(cls, path: str) -> "DeviceTree":
        out = C.c_void_p()
        capi.check(capi.hip().hm_hip_tree_load(path.encode(), C.byref(out), None), "tree_load")
        return cls(out)

    def multiply(self, other: "DeviceTree", thr: float) -> "DeviceTree":
        """self @ other as a new recompressed device tree (H x H)."""
        out = C.c_void_p()
        capi.check(capi.hip().hm_hip_multiply_hodlr_hodlr(self.h, other.h, thr, C.byref(out),
                                                          None), "multiply_hodlr_hodlr")
        return DeviceTree(out)

    FORMATS = {"fp64": 0, "bf16": 1, "int8": 2, "fp32": 3}

    def set_precision(self, slices: int, fmt: str = "bf16"):
        """Off-diagonal products in FP64 (slices 0, or fmt "fp64"), on 1..4
        BF16 or INT8 slices of the factors, or on one FP32 copy (fmt "fp32",
        slices 1)."""
        if slices == 0 or fmt == "fp64":
            code, slices = 0, 0
        else:
            code = self.FORMATS[fmt]
        capi.check(capi.hip().hm_hip_tree_set_format(self.h, code, slices, None),
                   "set_precision")

    def download(self, lib, tree_ptr):
        capi.check(capi.hip().hm_hip_tree_download(self.h, tree_ptr, None, None),
                   "tree download")

    def free(self):
        if self.h:
            capi.hip().hm_hip_tree_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def recompress_batch(us, vs, thr, method=-1):
    """Recompress factor pairs (numpy) on the device; returns new (U, V) lists."""
    n = len(us)
    ms = (C.c_int * n)(*[u.shape[0] for u in us])
    ns = (C.c_int * n)(*[v.shape[0] for v in vs])
    ss = (C.c_int * n)(*[u.shape[1] for u in us])
    du = [DeviceArray.from_numpy(u) for u in us]
    dv = [DeviceArray.from_numpy(v) for v in vs]
    pu = (C.c_void_p * n)(*[d.ptr.value for d in du])
    pv = (C.c_void_p * n)(*[d.ptr.value for d in dv])
    ks = (C.c_int * n)()
    capi.check(capi.hip().hm_hip_recompress_batch(n, ms, ns, ss, pu, pv, thr, method, ks,
                                                  None), "recompress_batch")
    outu, outv = [], []
    for i in range(n):
        k = ks[i]
        U = du[i].numpy()[:, :k].copy()
        V = dv[i].numpy()[:, :k].copy()
        outu.append(U)
        outv.append(V)
    return outu, outv


def compress_batch(mats, heights, thr, opts: capi.CompressOpts | None = None):
    """Compress device matrices (DeviceArray, square) in one batched call;
    returns (list of DeviceTree, stats)."""
    n = len(mats)
    pa = (C.c_void_p * n)(*[a.ptr.value for a in mats])
    ld = (C.c_long * n)(*[a.shape[0] for a in mats])
    ms = (C.c_int * n)(*[a.shape[0] for a in mats])
    hs = (C.c_int * n)(*[int(h) for h in heights])
    out = (C.c_void_p * n)()
    st = capi.CompressStats()
    capi.check(capi.hip().hm_hip_compress_batch(
        n, pa, ld, ms, hs, thr, C.byref(opts) if opts is not None else None, out,
        C.byref(st), None), "compress_batch")
    trees = []
    for i in range(n):
        t = DeviceTree(C.c_void_p(out[i]))
        t.stats = st
        trees.append(t)
    return trees, st


class Batch:
    """A batched product over many trees with fixed operands (hm_batch*)."""

    def __init__(self, trees, xs, ys, trans=0):
        n = len(trees)
        self.trees, self.xs, self.ys = trees, xs, ys
        pt = (C.c_void_p * n)(*[t.h.value for t in trees])
        # trans 0 / 1: X, Y are n x R; trans 2 (Y = X H): R x n
        R = (C.c_int * n)(*[(x.shape[0] if trans == 2 else
                             (x.shape[1] if len(x.shape) > 1 else 1)) for x in xs])
        px = (C.c_void_p * n)(*[x.ptr.value for x in xs])
        py = (C.c_void_p * n)(*[y.ptr.value for y in ys])
        lx = (C.c_long * n)(*[x.shape[0] for x in xs])
        ly = (C.c_long * n)(*[y.shape[0] for y in ys])
        self.h = C.c_void_p()
        capi.check(capi.hip().hm_hip_batch_create(n, pt, trans, R, px, lx, py, ly,
                                                  C.byref(self.h)), "batch_create")

    def tune(self):
        capi.check(capi.hip().hm_hip_batch_tune(self.h, None), "batch_tune")

    def run(self, graph=False):
        capi.check(capi.hip().hm_hip_batch_run(self.h, 1 if graph else 0, None), "batch_run")

    def variants(self):
        buf = (C.c_int * 64)()
        n = capi.hip().hm_hip_batch_variants(self.h, buf, 64)
        return list(buf[:min(n, 64)])

    def free(self):
        if self.h:
            capi.hip().hm_hip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def kernel_timing(enable: bool, only: str | None = None):
    capi.hip().hm_hip_kernel_timing_select(only.encode() if only else None)
    capi.hip().hm_hip_kernel_timing(1 if enable else 0)


def kernel_timing_read(name: str):
    tot = C.c_double(0)
    cnt = C.c_long(0)
    capi.hip().hm_hip_kernel_timing_read(name.encode(), C.byref(tot), C.byref(cnt))
    return tot.value, cnt.value


def kernel_timing_dump():
    """{name: (total_ms, count)} of every timed kernel / host section."""
    n = capi.hip().hm_hip_kernel_timing_dump(None, 0)
    buf = C.create_string_buffer(n + 1)
    capi.hip().hm_hip_kernel_timing_dump(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        name, tot, cnt = line.split("\t")
        out[name] = (float(tot), int(cnt))
    return out


def kernel_timing_reset():
    capi.hip().hm_hip_kernel_timing_reset()
