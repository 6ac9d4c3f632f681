"""Thin QR of device panels (hm_hip_qr_batch): the TSQR chunk kernels and
the block Gram-Schmidt across 64-column panels, on single-chunk, multi-level
and ragged shapes batched together, vs numpy's Householder QR (LAPACK).

Checks: Y = Q R to 1e-13 relative, Q^T Q = I to 1e-13, R upper triangular
with exact zeros below the diagonal, and |R| equal to LAPACK's |R| row by
row to 1e-12 relative (a TSQR tree may flip the sign of a row of R and the
matching column of Q).  A zero column (tau = 0) and a repeated column (a
numerically rank-deficient single panel) are included."""
import ctypes as C
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (5, 3), (64, 64), (100, 17), (128, 64), (129, 64), (300, 33),
          (1000, 16), (4097, 64), (2000, 130), (777, 100), (512, 1)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _panel(rng, m, n, kind):
    y = rng.standard_normal((m, n))
    if kind == "zero_col" and n > 2:
        y[:, 1] = 0.0
    if kind == "repeat_col" and n > 3:
        y[:, 3] = y[:, 2]
    return np.asfortranarray(y)


def _run(panels, with_r=True):
    from hmat_lib_amd import device as D
    n_jobs = len(panels)
    dy, dq, dr = [], [], []
    for i, y in enumerate(panels):
        m, n = y.shape
        dy.append(D.DeviceArray.from_numpy(y))
        dq.append(D.DeviceArray((m, n), zero=True))
        dr.append(D.DeviceArray((n, n), zero=True) if with_r else None)
    V = C.c_void_p
    arr = lambda xs: (V * n_jobs)(*[x.ptr.value if x is not None else None for x in xs])
    lds = lambda xs: (C.c_long * n_jobs)(*[x.shape[0] if x is not None else 1 for x in xs])
    ms = (C.c_int * n_jobs)(*[y.shape[0] for y in panels])
    ns = (C.c_int * n_jobs)(*[y.shape[1] for y in panels])
    rc = capi.hip().hm_hip_qr_batch(n_jobs, ms, ns, arr(dy), lds(dy), arr(dq), lds(dq),
                                    arr(dr) if with_r else None, lds(dr) if with_r else None,
                                    None)
    capi.check(rc, "hm_hip_qr_batch")
    D.synchronize()
    out = [(q.numpy(), r.numpy() if r is not None else None) for q, r in zip(dq, dr)]
    for a in dy + dq + [r for r in dr if r is not None]:
        a.free()
    return out


def _check(y, q, r, label):
    m, n = y.shape
    ny = max(np.linalg.norm(y), 1e-300)
    assert np.linalg.norm(q @ r - y) <= 1e-13 * ny, label
    assert np.linalg.norm(q.T @ q - np.eye(n)) <= 1e-13 * max(1.0, np.sqrt(n)), label
    assert np.all(np.tril(r, -1) == 0.0), label
    r_np = np.linalg.qr(y, mode="r")
    nr = max(np.linalg.norm(r_np), 1e-300)
    assert np.linalg.norm(np.abs(r) - np.abs(r_np)) <= 1e-12 * nr, label


def test_qr_batch_shapes(gpu):
    rng = np.random.default_rng(11)
    panels = [_panel(rng, m, n, "plain") for m, n in SHAPES]
    _say(f"qr batch of {len(panels)} panels")
    for (q, r), y, s in zip(_run(panels), panels, SHAPES):
        _check(y, q, r, f"shape {s}")


@pytest.mark.parametrize("kind", ["zero_col", "repeat_col"])
def test_qr_degenerate_columns(gpu, kind):
    rng = np.random.default_rng(12)
    # (3000, 16) and (1000, 12): narrow panels over many chunks and levels of
    # the one-wave TSQR kernels (panels of at most 16 columns)
    shapes = [(64, 8), (300, 40), (3000, 64), (3000, 16), (1000, 12)]
    panels = [_panel(rng, m, n, kind) for m, n in shapes]
    _say(f"qr {kind}")
    for (q, r), y, s in zip(_run(panels), panels, shapes):
        m, n = y.shape
        assert np.linalg.norm(q @ r - y) <= 1e-13 * np.linalg.norm(y), s
        assert np.all(np.tril(r, -1) == 0.0), s
        # the dependent column leaves a zero (or rounding-level) pivot; the
        # reflector of a zero column is the identity (tau = 0), so Q stays
        # orthonormal either way
        assert np.linalg.norm(q.T @ q - np.eye(n)) <= 1e-12 * np.sqrt(n), s
        col = 1 if kind == "zero_col" else 3
        assert abs(r[col, col]) <= 1e-12 * np.linalg.norm(y), s


def test_qr_without_r(gpu):
    rng = np.random.default_rng(13)
    panels = [_panel(rng, 700, 70, "plain")]
    (q, _), = _run(panels, with_r=False)
    y = panels[0]
    assert np.linalg.norm(q.T @ q - np.eye(70)) <= 1e-13 * np.sqrt(70)
    assert np.linalg.norm(q @ (q.T @ y) - y) <= 1e-13 * np.linalg.norm(y)


def test_qr_rejects_bad_shapes(gpu):
    from hmat_lib_amd import device as D
    y = D.DeviceArray((3, 5), zero=True)
    q = D.DeviceArray((3, 5), zero=True)
    V = C.c_void_p
    rc = capi.hip().hm_hip_qr_batch(1, (C.c_int * 1)(3), (C.c_int * 1)(5), (V * 1)(y.ptr.value),
                                    (C.c_long * 1)(3), (V * 1)(q.ptr.value), (C.c_long * 1)(3),
                                    None, None, None)
    assert rc != 0
    y.free()
    q.free()
