"""Learned rank/sketch predictor for the randomized compression.

A small MLP (hm_hip_mlp_*: FP32, nin -> 64 -> 64 -> 1, tanh) predicts, for
an off-diagonal block, log2(w + 1) where w is the number of its singular
values at or above accept/20 * thr * sigma_0 (the rank at a threshold 200x
tighter): with 8 columns of oversampling a sketch of that width leaves a
residual below the test's accept * thr * sigma_0 - from
  * 7 geometric inputs: log2 of the block's row and column counts, its
    level, the distance between the centroids of its row and column points,
    the kernel length scale, and the bounding-box diagonals of the two point
    sets (all log-scaled), and
  * the 16 singular values of the block's first, 16-column sketch (the
    library appends 17 features derived from them, hm_hip_rank_features).
Blocks whose 16-column sketch fails the residual test get a second sketch
of ceil(w (1 + margin)) + 8 columns (w predicted) instead of the
doubling sequence 32, 64, ...; the residual test still decides, and a
failed prediction goes on doubling (the plain path), so a bad model costs
time, never accuracy.

Training data are the widths the plain path itself finds on seeded
kernel matrices (labels, hm_compress_opts.need_out), with the first-round
spectra it records (hm_compress_opts.first_sigma).  Training runs on the GPU through the
library's forward/backward/Adam kernels.  `python -m hmat_lib_amd.learned`
regenerates the packaged weights (data/rank_mlp.npz).
"""
from __future__ import annotations

import ctypes as C
import os
import time

import numpy as np

from . import capi
from . import workloads as WL

N_EXTRA = 7
N_HIDDEN = 64
N_IN = N_EXTRA + 17
SKETCH_MAX = 1024  # widest sketch, plain and learned alike
HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "data", "rank_mlp.npz")
FP = C.POINTER(C.c_float)


def _fptr(a):
    return a.ctypes.data_as(FP)


def leaf_offsets(n, height):
    """Leaf offsets by halving (top-left gets the larger half)."""
    sizes = [n]
    for _ in range(height):
        sizes = [s for m in sizes for s in (m - m // 2, m // 2)]
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)


def block_features(pts, height, ell):
    """(nblock, 7) geometric inputs, level-major blocks (level 1 first).
    Every tree node is a contiguous range of the points: per-leaf sums and
    bounding boxes, combined up the tree."""
    pts = np.asarray(pts, dtype=np.float64)
    if pts.ndim == 1:
        pts = pts[:, None]
    n, d = pts.shape
    off = leaf_offsets(n, height)
    nl = 1 << height
    cnt = np.diff(off).astype(np.int64)
    full = cnt > 0
    lsum = np.zeros((nl, d))
    llo = np.full((nl, d), np.inf)
    lhi = np.full((nl, d), -np.inf)
    if full.any():
        starts = off[:-1][full]
        lsum[full] = np.add.reduceat(pts, starts, axis=0)
        llo[full] = np.minimum.reduceat(pts, starts, axis=0)
        lhi[full] = np.maximum.reduceat(pts, starts, axis=0)
    out = []
    for lvl in range(1, height + 1):
        nn, span = 1 << lvl, 1 << (height - lvl)
        c = cnt.reshape(nn, span).sum(1)
        sm = lsum.reshape(nn, span, d).sum(1)
        lo = llo.reshape(nn, span, d).min(1)
        hi = lhi.reshape(nn, span, d).max(1)
        cen = np.where(c[:, None] > 0, sm / np.maximum(c, 1)[:, None], 0.0)
        dia = np.where(c > 0, np.linalg.norm(np.where(c[:, None] > 0, hi - lo, 0.0), axis=1), 0.0)
        part = np.arange(nn) ^ 1
        m, k = np.maximum(c, 1), np.maximum(c[part], 1)
        sep = np.linalg.norm(cen - cen[part], axis=1)
        out.append(np.stack([np.log2(m) / 16.0, np.log2(k) / 16.0, np.full(nn, lvl / 16.0),
                             np.log10(sep + 1e-3) / 3.0, np.full(nn, np.log10(ell) / 2.0),
                             np.log10(dia + 1e-6) / 3.0, np.log10(dia[part] + 1e-6) / 3.0],
                            axis=1))
    return np.concatenate(out).astype(np.float32)


def rank_features(extra, sig16, thr):
    """Model inputs (library definition): extra features + 17 spectral ones."""
    extra = np.ascontiguousarray(extra, dtype=np.float32)
    sig16 = np.ascontiguousarray(sig16, dtype=np.float64)
    nb = sig16.shape[0]
    out = np.zeros((nb, extra.shape[1] + 17), dtype=np.float32)
    capi.check(capi.hip().hm_hip_rank_features(nb, extra.shape[1], _fptr(extra),
                                               capi.dptr(sig16), thr, _fptr(out)),
               "rank_features")
    return out


class RankModel:
    """hm_mlp* owned by Python."""

    def __init__(self, nin=N_IN, nh=N_HIDDEN, seed=1):
        self.nin, self.nh = nin, nh
        self.h = C.c_void_p()
        capi.check(capi.hip().hm_hip_mlp_create(nin, nh, seed, C.byref(self.h)), "mlp_create")

    @property
    def nparams(self):
        return int(capi.hip().hm_hip_mlp_nparams(self.h))

    def params(self):
        p = np.zeros(self.nparams, dtype=np.float32)
        capi.check(capi.hip().hm_hip_mlp_get_params(self.h, _fptr(p)), "mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.nparams
        capi.check(capi.hip().hm_hip_mlp_set_params(self.h, _fptr(p)), "mlp_set_params")

    def _dev(self, a):
        from . import device as D
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = D.DeviceArray((max(1, (a.size + 1) // 2),))  # float storage in a double buffer
        capi.check(capi.hip().hm_hip_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes,
                                                None), "h2d")
        return d

    def forward(self, X):
        from . import device as D
        X = np.ascontiguousarray(X, dtype=np.float32)
        dx = self._dev(X)
        dy = D.DeviceArray((max(1, (X.shape[0] + 1) // 2),))
        capi.check(capi.hip().hm_hip_mlp_forward(self.h, dx.ptr, X.shape[0], dy.ptr, None),
                   "mlp_forward")
        y = np.zeros(X.shape[0], dtype=np.float32)
        capi.check(capi.hip().hm_hip_memcpy_d2h(y.ctypes.data_as(C.c_void_p), dy.ptr, y.nbytes,
                                                None), "d2h")
        D.synchronize()
        return y

    def gradient(self, X, t):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = self._dev(X), self._dev(t)
        g = np.zeros(self.nparams, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        capi.check(capi.hip().hm_hip_mlp_gradient(self.h, dx.ptr, dt.ptr, X.shape[0], _fptr(g),
                                                  _fptr(loss), None), "mlp_gradient")
        return float(loss[0]), g

    def train(self, X, t, epochs=200, batch=256, lr=2e-3, seed=7):
        X = np.ascontiguousarray(X, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = self._dev(X), self._dev(t)
        loss = np.zeros(epochs, dtype=np.float32)
        capi.check(capi.hip().hm_hip_mlp_train(self.h, dx.ptr, dt.ptr, X.shape[0], epochs, batch,
                                               lr, seed, _fptr(loss), None), "mlp_train")
        return loss

    def save(self, path=WEIGHTS, **meta):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.params(), nin=self.nin, nh=self.nh, **meta)

    @classmethod
    def load(cls, path=WEIGHTS):
        z = np.load(path, allow_pickle=False)
        m = cls(int(z["nin"]), int(z["nh"]))
        m.set_params(z["params"])
        return m

    def free(self):
        if self.h:
            capi.hip().hm_hip_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def compress(da, pts, height, thr, ell, model: RankModel | None, opts=None):
    """Compress device matrix `da` (n x n over points pts) with the learned
    second width (model) or the plain doubling (model None).  Returns the
    DeviceTree and the first-round spectra (nblock x 16)."""
    from . import device as D
    n = da.shape[0]
    if opts is None:
        opts = D.default_compress_opts()
        opts.sketch_max = SKETCH_MAX
    o = opts
    nblock = (1 << (height + 1)) - 2
    sig = np.zeros((nblock, 16), dtype=np.float64)
    need = np.zeros(nblock, dtype=np.int32)
    o.first_sigma = capi.dptr(sig)
    o.need_out = need.ctypes.data_as(capi.PInt)
    feats = None
    if model is not None:
        feats = np.ascontiguousarray(block_features(pts, height, ell), dtype=np.float32)
        o.rank_model = model.h
        o.block_features = _fptr(feats)
        o.n_extra = N_EXTRA
    t = D.DeviceTree.compress(da, n, height, thr, opts=o)
    o.first_sigma = None
    o.need_out = None
    o.rank_model = None
    o.block_features = None
    t.need = need
    return t, sig


def make_dataset(instances, thr=1e-8, kind=2, log=print):
    """Plain compressions of seeded kernel matrices -> (X, y, info).
    instances: iterable of (n, height, ell, seed)."""
    from . import device as D
    Xs, ys = [], []
    for (n, height, ell, seed) in instances:
        pts = WL.points_2d_morton(seed, n)
        da = D.fill_kernel_matrix(kind, pts, ell, 0.0)
        t0 = time.perf_counter()
        tr, sig = compress(da, pts, height, thr, ell, None)
        ranks = tr.ranks()
        X = rank_features(block_features(pts, height, ell), sig, thr)
        Xs.append(X)
        ys.append(np.log2(tr.need + 1.0).astype(np.float32))
        if len(Xs) % 50 == 1:
            log(f"[learned] {len(Xs)}: n={n} h={height} ell={ell:.4f} seed={seed}: "
                f"{(time.perf_counter() - t0) * 1e3:.0f} ms, max rank {ranks.max()}, "
                f"max width {tr.need.max()}")
        tr.free()
        da.free()
    return np.vstack(Xs), np.concatenate(ys)


def training_instances(count, seed=5005):
    """Seeded Matern-3/2 instances: ell ~ U[0.01, 1], leaf 128, N = 4096,
    8192, 16384 or 32768 with probabilities 0.35, 0.35, 0.2, 0.1 (the
    blocks of every level of a 32768 tree occur in the smaller trees too,
    except the two largest levels)."""
    u = WL.uniform(seed, 2 * count).reshape(count, 2)
    out = []
    for i in range(count):
        v = float(u[i, 1])
        n = 4096 if v < 0.35 else (8192 if v < 0.7 else (16384 if v < 0.9 else 32768))
        h = int(np.log2(n // 128))
        out.append((n, h, 0.01 + 0.99 * float(u[i, 0]), seed + 11 * i))
    return out


def main():
    import argparse
    from . import device as D
    ap = argparse.ArgumentParser(description="train the packaged rank predictor on the GPU")
    ap.add_argument("--instances", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--dataset", default=None, help="also save the dataset (npz)")
    args = ap.parse_args()
    D.init(0)
    inst = training_instances(args.instances)
    X, y = make_dataset(inst)
    if args.dataset:
        np.savez(args.dataset, X=X, y=y)
    model = RankModel()
    t0 = time.perf_counter()
    loss = model.train(X, y, epochs=args.epochs)
    pred = model.forward(X)
    k, kp = np.exp2(y) - 1, np.exp2(pred) - 1
    print(f"[learned] trained on {len(y)} blocks in {time.perf_counter() - t0:.1f} s; "
          f"loss {loss[0]:.4f} -> {loss[-1]:.4f}; median |k_pred/k - 1| "
          f"{np.median(np.abs(kp / np.maximum(k, 1) - 1)):.3f}", flush=True)
    model.save(args.out, instances=np.array(inst, dtype=np.float64), loss=loss)


if __name__ == "__main__":
    main()
