"""Synthetic inputs of the HODLR workloads (seeded, reproducible).

RNG: splitmix64; uniform u = (z >> 11) * 2^-53; normals by Box-Muller on
pairs (u1, u2): sqrt(-2 ln(1 - u1)) cos(2 pi u2).  Points are generated on
the host and uploaded; kernel matrices are generated on the device.

Workloads (matrix order N, tree height h, leaf = N / 2^h):
  exp1d_1k   N=1024, x ~ U[0,1] sorted, exp(-|x-y|/0.1) + 1e-3 I, h=3,
             thr 1e-8, x ~ N(0,1), X 1024 x 16
  exp1d_64k  N=65536, sorted uniform, exp(-|x-y|/0.05), h=8 (leaf 256),
             thr 1e-10, X 65536 x 64
  inv2d_64k  N=65536, U[0,1]^2 Morton-ordered, 1/(|x-y| + 1e-2), h=9
             (leaf 128), thr 1e-8, X 65536 x 256
  matern_32k N=32768, U[0,1]^2 Morton, Matern-3/2 with ell, h=8, thr 1e-8
  batch_512  512 exp1d matrices, N log-uniform in [512, 8192] rounded to a
             multiple of 64, ell ~ U[0.02, 0.5], leaf 128, thr 1e-8,
             RHS log-uniform in [1, 128]
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """count outputs of splitmix64 started at `seed`, skipping `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, count, offset) >> np.uint64(11)).astype(np.float64) * 2.0 ** -53


def normal(seed: int, count: int) -> np.ndarray:
    u = uniform(seed, 2 * count)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def normal_matrix(seed: int, rows: int, cols: int) -> np.ndarray:
    return np.asfortranarray(normal(seed, rows * cols).reshape(rows, cols, order="F"))


def points_1d_sorted(seed: int, n: int) -> np.ndarray:
    return np.sort(uniform(seed, n))


def morton_order(pts: np.ndarray) -> np.ndarray:
    """Indices sorting 2-D points in [0,1]^2 by 16-bit-per-axis Morton code
    (ties broken by index)."""
    q = np.clip((pts * 65536.0).astype(np.int64), 0, 65535).astype(np.uint64)

    def spread(x):
        x = (x | (x << np.uint64(8))) & np.uint64(0x00FF00FF)
        x = (x | (x << np.uint64(4))) & np.uint64(0x0F0F0F0F)
        x = (x | (x << np.uint64(2))) & np.uint64(0x33333333)
        x = (x | (x << np.uint64(1))) & np.uint64(0x55555555)
        return x

    code = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1))
    return np.lexsort((np.arange(len(code)), code))


def points_2d_morton(seed: int, n: int) -> np.ndarray:
    u = uniform(seed, 2 * n).reshape(n, 2)
    return np.ascontiguousarray(u[morton_order(u)])


@dataclass
class Workload:
    name: str
    n: int
    height: int
    kind: int          # 0 exp, 1 inverse distance, 2 Matern-3/2
    ell: float
    shift: float
    thr: float
    rhs: int
    dim: int
    seed: int

    @property
    def leaf(self):
        return self.n >> self.height

    def points(self):
        if self.dim == 1:
            return points_1d_sorted(self.seed, self.n)
        return points_2d_morton(self.seed, self.n)

    def rhs_matrix(self, cols=None):
        return normal_matrix(self.seed + 1, self.n, cols or self.rhs)


WORKLOADS = {
    "exp1d_1k": Workload("exp1d_1k", 1024, 3, 0, 0.1, 1e-3, 1e-8, 16, 1, 1001),
    "exp1d_64k": Workload("exp1d_64k", 65536, 8, 0, 0.05, 0.0, 1e-10, 64, 1, 2002),
    "inv2d_64k": Workload("inv2d_64k", 65536, 9, 1, 1e-2, 0.0, 1e-8, 256, 2, 4004),
    "matern_32k": Workload("matern_32k", 32768, 8, 2, 0.1, 0.0, 1e-8, 16, 2, 5005),
}


def batch_workloads(count=512, seed=3003):
    """The uneven batch: (n, ell, rhs, height) per matrix."""
    u = uniform(seed, 3 * count).reshape(count, 3)
    n = np.exp(np.log(512) + u[:, 0] * (np.log(8192) - np.log(512)))
    n = np.maximum(64, (np.rint(n / 64) * 64).astype(np.int64))
    ell = 0.02 + u[:, 1] * 0.48
    rhs = np.rint(np.exp(u[:, 2] * np.log(128))).astype(np.int64)
    height = np.maximum(1, np.rint(np.log2(n / 128)).astype(np.int64))
    return [dict(n=int(a), ell=float(b), rhs=int(c), height=int(d), seed=seed + 7 * i)
            for i, (a, b, c, d) in enumerate(zip(n, ell, rhs, height))]
