"""The learned rank/sketch predictor: the MLP's own GPU forward, backward and
Adam kernels vs the FP64 oracle (oracle/mlp_oracle.py), deterministic
training, and what the predictor changes in compression: fewer sketch
rounds at the same ranks and accuracy, with the residual test catching a
wrong prediction (fallback to the doubling path).

Tolerances: FP32 network vs FP64 oracle 2e-4 relative (forward) and 1e-3
relative in norm (gradient); compression as in test_gpu_compress."""
import os
import sys

import numpy as np
import pytest

from hmat_lib_amd import workloads as WL
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mlp_forward_gradient_vs_oracle(gpu):
    from hmat_lib_amd import learned as L
    from oracle import mlp_oracle as M
    _say("case: mlp forward/gradient vs oracle")
    rng = np.random.default_rng(21)
    m = L.RankModel(seed=3)
    p = m.params()
    assert p.size == M.nparams(L.N_IN, L.N_HIDDEN)
    X = rng.standard_normal((300, L.N_IN)).astype(np.float32)
    t = rng.standard_normal(300).astype(np.float32)
    y = m.forward(X)
    yo, _, _ = M.forward(p.astype(np.float64), X.astype(np.float64), L.N_HIDDEN)
    np.testing.assert_allclose(y, yo, rtol=2e-4, atol=2e-5)
    loss, g = m.gradient(X, t)
    lo, go = M.loss_and_grad(p.astype(np.float64), X.astype(np.float64), t.astype(np.float64),
                             L.N_HIDDEN)
    assert abs(loss - lo) <= 1e-4 * lo
    assert np.linalg.norm(g - go) <= 1e-3 * np.linalg.norm(go)
    # every parameter block carries gradient (no block silently zero)
    o = 0
    for size in (L.N_IN * 64, 64, 64 * 64, 64, 64, 1):
        assert np.linalg.norm(g[o:o + size]) > 0
        o += size
    m.free()


def test_mlp_training_deterministic_and_learns(gpu):
    from hmat_lib_amd import learned as L
    _say("case: mlp training")
    rng = np.random.default_rng(22)
    X = rng.uniform(-1, 1, size=(2048, L.N_IN)).astype(np.float32)
    t = (np.sin(2 * X[:, 0]) + X[:, 1] * X[:, 2] + 0.5 * X[:, 3]).astype(np.float32)
    runs = []
    for _ in range(2):
        m = L.RankModel(seed=9)
        loss = m.train(X, t, epochs=60, batch=128, lr=3e-3, seed=4)
        runs.append((loss, m.params()))
        m.free()
    assert runs[0][0][-1] < 0.1 * runs[0][0][0]
    np.testing.assert_array_equal(runs[0][1], runs[1][1])


def _matern(n, ell, seed):
    from hmat_lib_amd import device as D
    pts = WL.points_2d_morton(seed, n)
    return pts, D.fill_kernel_matrix(2, pts, ell, 0.0)


def _product_err(tree, pts, n, ell, rows=96):
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(1)
    x = np.asfortranarray(rng.standard_normal((n, 4)))
    dx, dy = D.DeviceArray.from_numpy(x), D.DeviceArray((n, 4))
    tree.matmat(dx, dy)
    y = dy.numpy()
    r = np.sort(rng.choice(n, rows, replace=False))
    d = np.sqrt(((pts[r][:, None, :] - pts[None, :, :]) ** 2).sum(-1))
    tt = np.sqrt(3.0) * d / ell
    ref = ((1 + tt) * np.exp(-tt)) @ x
    return np.linalg.norm(y[r] - ref) / np.linalg.norm(ref)


def test_bad_predictor_falls_back(gpu):
    """A model that predicts rank 0 everywhere: every widened block fails its
    residual test and continues on the doubling path, so ranks equal the
    plain path's and accuracy is unchanged."""
    from hmat_lib_amd import learned as L
    _say("case: bad predictor fallback")
    n, h, ell = 4096, 5, 0.2
    pts, da = _matern(n, ell, 61)
    m = L.RankModel()
    m.set_params(np.zeros(m.nparams, dtype=np.float32))
    plain, _ = L.compress(da, pts, h, 1e-8, ell, None)
    bad, _ = L.compress(da, pts, h, 1e-8, ell, m)
    st = bad.stats
    assert st.predicted_blocks > 0 and st.predicted_failed == st.predicted_blocks
    np.testing.assert_array_equal(bad.ranks(), plain.ranks())
    assert _product_err(bad, pts, n, ell) < 1e-7
    m.free()


@pytest.mark.skipif(not os.path.exists(os.path.join(
    os.path.dirname(__file__), "..", "hmat_lib_amd", "data", "rank_mlp.npz")),
    reason="packaged weights not built")
def test_trained_predictor_gains(gpu):
    """The packaged model on held-out Matern instances (length scales and
    point seeds not in its training set): the predicted second width
    replaces the doubling rounds, ranks agree with the plain path (a block
    may move by one where a singular value sits at the threshold), the
    product error is unchanged, most predictions pass the residual test, and
    the compression takes less wall time than the plain path's (best of
    three runs each)."""
    from hmat_lib_amd import learned as L
    from hmat_lib_amd import device as D
    import time
    model = L.RankModel.load()
    t_plain = t_learned = 0.0
    for (n, h, ell, seed) in [(16384, 7, 0.04, 777), (16384, 7, 0.45, 778)]:
        _say(f"case: trained predictor n={n} ell={ell}")
        pts, da = _matern(n, ell, seed)
        best = {}
        for mode in ("plain", "learned", "plain", "learned", "plain", "learned"):
            D.synchronize()
            t0 = time.perf_counter()
            tr, _ = L.compress(da, pts, h, 1e-8, ell, model if mode == "learned" else None)
            D.synchronize()
            el = time.perf_counter() - t0
            if mode not in best or el < best[mode][0]:
                if mode in best:
                    best[mode][1].free()
                best[mode] = (el, tr)
            else:
                tr.free()
        (tp, plain), (tl, learned) = best["plain"], best["learned"]
        t_plain += tp
        t_learned += tl
        sp, sl = plain.stats, learned.stats
        _say(f"  plain {sp.ms_total:.0f} ms rounds {sp.sketch_rounds}; learned "
             f"{sl.ms_total:.0f} ms rounds {sl.sketch_rounds}, predicted "
             f"{sl.predicted_blocks} failed {sl.predicted_failed}")
        rp, rl = plain.ranks(), learned.ranks()
        assert np.all(np.abs(rp - rl) <= 1)
        assert (rp != rl).sum() <= max(2, len(rp) // 50)
        assert sl.sketch_rounds <= sp.sketch_rounds
        assert sl.predicted_failed <= 0.25 * max(1, sl.predicted_blocks)
        ep, el = _product_err(plain, pts, n, ell), _product_err(learned, pts, n, ell)
        assert el < max(2 * ep, 1e-8)
        plain.free(), learned.free(), da.free()
    model.free()
    # the predictor saves sketch rounds, and with them time (best of three
    # compressions each, over both matrices)
    _say(f"  wall time plain {t_plain * 1e3:.0f} ms, learned {t_learned * 1e3:.0f} ms")
    assert t_learned < 0.95 * t_plain
