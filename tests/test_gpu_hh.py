"""HODLR x HODLR product on the GPU (hm_hip_multiply_hodlr_hodlr and the
drop-in multiply_hodlr_hodlr) vs the reference's known answers, the oracle
restatement and exact algebra.

Tolerances: the reference's DELTA, 1e-10 absolute, on reconstructions of the
exactly low-rank known answers (tests/utils/expect.h:9); blocks vs the
oracle's product: ranks equal (or within one where a singular value sits
near the threshold) and |U V^T - U_o V_o^T|_2 <= 4 thr (both truncate at
the absolute threshold thr); products of the result vs the dense product
1e-8 relative."""
import ctypes as C
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from hmat_lib_amd import workloads as WL
from tests import helpers as H
from tests.test_gpu_compress import MS155, _dropin_compress

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _dropin_hh(lib, ta, tb, height, thr):
    ierr = C.c_int(-1)
    tc = lib.allocate_tree_monolithic(height, C.byref(ierr))
    assert ierr.value == 0
    info = lib.multiply_hodlr_hodlr(ta, tb, tc, thr, C.byref(ierr))
    assert ierr.value == 0 and info == 0, capi.last_error()
    return tc


def _compare_blocks(fc, fo, thr, max_mismatch=0):
    mism = 0
    for key, (u, v) in fc.blocks.items():
        uo, vo = fo.blocks[key]
        mism += u.shape[1] != uo.shape[1]
        assert abs(u.shape[1] - uo.shape[1]) <= 1, (key, u.shape[1], uo.shape[1])
        d = u @ v.T - uo @ vo.T
        assert np.linalg.norm(d, 2) <= 4 * thr + 1e-13 * np.linalg.norm(uo @ vo.T, 2), key
    assert mism <= max_mismatch, mism
    for i, d in enumerate(fc.diag):
        np.testing.assert_allclose(d, fo.diag[i], rtol=1e-12, atol=1e-13 * max(1.0, np.abs(d).max()))


@pytest.mark.parametrize("height", [1, 2, 3])
def test_known_answers(gpu, height):
    """test_hodlr_hodlr_algebra.c:100-339 through the drop-in: Laplacian x
    converse (plain and with corners), identity x {I, 0, L}, and I, 0, L
    times themselves (same tree as both operands)."""
    for name, a, b, exp, same in H.hh_known_answers():
        _say(f"case: hh {name} h={height}")
        lib, ta = _dropin_compress(a, height, 1e-8)
        tb = ta if same else _dropin_compress(b, height, 1e-8)[1]
        tc = _dropin_hh(lib, ta, tb, height, 1e-8)
        fc = T.flatten(tc)
        np.testing.assert_allclose(fc.to_dense(), exp, atol=1e-10, err_msg=name)
        oa = H.oracle_compress(a, height, 1e-8)
        ob = oa if same else H.oracle_compress(b, height, 1e-8)
        oc = H.oracle_hh(oa, ob, 1e-8)
        assert fc.ranks() == T.flatten(oc).ranks(), name
        for t in {id(x): x for x in (oa, ob, oc)}.values():
            H.oracle_free(t)
        for t in {id(x): x for x in (ta, tb, tc)}.values():
            lib.free_tree_hodlr(C.byref(t))


def test_decay_products_vs_oracle(gpu):
    """glibc decay matrices (test_constructors.c:1047-1105, m = 155),
    heights 1-5, halves and custom leaf sizes: every block of the product
    vs the oracle's (whose agreement with the reference is bitwise)."""
    mats = H.glibc_decay_matrices(155, 1.0, True, 6, 6)
    k = 0
    for height in range(1, 6):
        for custom in (False, True):
            a, b = mats[k % 12][1], mats[(k + 5) % 12][1]
            k += 1
            ms = MS155[height] if custom else None
            _say(f"case: hh decay h={height} custom={custom}")
            lib, ta = _dropin_compress(a, height, 1e-8, ms)
            _, tb = _dropin_compress(b, height, 1e-8, ms)
            oa, ob = H.oracle_compress(a, height, 1e-8, ms), H.oracle_compress(b, height, 1e-8, ms)
            for same in (False, True):
                tc = _dropin_hh(lib, ta, ta if same else tb, height, 1e-8)
                oc = H.oracle_hh(oa, oa if same else ob, 1e-8)
                _compare_blocks(T.flatten(tc), T.flatten(oc), 1e-8, max_mismatch=1)
                H.oracle_free(oc)
                lib.free_tree_hodlr(C.byref(tc))
            H.oracle_free(oa)
            H.oracle_free(ob)
            lib.free_tree_hodlr(C.byref(ta))
            lib.free_tree_hodlr(C.byref(tb))


def test_random_trees_exact_algebra(gpu):
    """Random trees with uneven leaves and ranks up to 12 (blocks whose
    summed rank reaches min(m, n) take the dense recompression): with a
    threshold far below the kept singular values the product equals the
    dense product."""
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(21)
    lib = capi.hmat()
    for height, leaves in ((2, [9, 14, 6, 11]), (3, [7, 12, 5, 9, 13, 8, 6, 10]),
                           (4, list(rng.integers(6, 30, size=16)))):
        _say(f"case: hh random h={height}")
        keep = T.Keep()
        fa = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(1, 6)))
        fb = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(1, 13)))
        ta, tb = T.fill_tree(lib, fa, keep), T.fill_tree(lib, fb, keep)
        da = D.DeviceTree.upload(ta.contents.root, height)
        db = D.DeviceTree.upload(tb.contents.root, height)
        dc = da.multiply(db, 1e-12)
        ierr = C.c_int(0)
        tc = lib.allocate_tree_monolithic(height, C.byref(ierr))
        dc.download(lib, tc)
        exp = fa.to_dense() @ fb.to_dense()
        assert H.rel_err(T.flatten(tc).to_dense(), exp) < 1e-12
        oc = H.oracle_hh(ta, tb, 1e-12)
        assert T.flatten(tc).ranks() == T.flatten(oc).ranks()
        H.oracle_free(oc)
        lib.free_tree_hodlr(C.byref(tc))
        for t in (da, db, dc):
            t.free()
        T.detach_and_free(lib, ta)
        T.detach_and_free(lib, tb)


@pytest.mark.parametrize("kind,n,height,ell,thr", [(0, 4096, 5, 0.05, 1e-10),
                                                   (1, 2048, 4, 1e-2, 1e-8)])
def test_kernel_products(gpu, kind, n, height, ell, thr):
    """Device-compressed kernel matrices (1-D exponential: ranks 1; 2-D
    inverse distance: ranks to ~40, QR recompression of s_tot up to a few
    hundred): A B through the device API, its products vs A (B X) of the
    dense matrices, and its ranks vs the oracle's product of the same two
    trees (downloaded).  The product's threshold is absolute (the
    reference's rule), so it is set to 1e-9 |A B X| / |X|: blocks of A B
    reach sigma_0 ~ 1e7 here, and an absolute 1e-8 would cut inside the
    FP64 rounding floor, where any two SVDs disagree on the rank."""
    from hmat_lib_amd import device as D
    _say(f"case: hh kernel kind={kind} n={n} h={height}")
    if kind == 0:
        pa = np.sort(WL.uniform(5, n))[:, None]
        pb = np.sort(WL.uniform(6, n))[:, None]
    else:
        pa = WL.points_2d_morton(77, n)
        pb = WL.points_2d_morton(78, n)
    daa = D.fill_kernel_matrix(kind, pa, ell, 0.0)
    dbb = D.fill_kernel_matrix(kind, pb, ell, 0.0)
    a, b = daa.numpy(), dbb.numpy()
    ta = D.DeviceTree.compress(daa, n, height, thr)
    tb = D.DeviceTree.compress(dbb, n, height, thr)
    x = np.asfortranarray(np.random.default_rng(2).standard_normal((n, 6)))
    ref = a @ (b @ x)
    thr_ab = 1e-9 * np.linalg.norm(ref) / np.linalg.norm(x)
    tc = ta.multiply(tb, thr_ab)
    dx = D.DeviceArray.from_numpy(x)
    dy = D.DeviceArray((n, 6))
    tc.matmat(dx, dy)
    assert H.rel_err(dy.numpy(), ref) < 1e-7
    lib = capi.hmat()
    ierr = C.c_int(0)
    ha = lib.allocate_tree_monolithic(height, C.byref(ierr))
    hb = lib.allocate_tree_monolithic(height, C.byref(ierr))
    ta.download(lib, ha)
    tb.download(lib, hb)
    oc = H.oracle_hh(ha, hb, thr_ab)
    ro = T.flatten(oc).ranks()
    rc = tc.ranks()
    keys = sorted(ro, key=lambda k: (1 << k[0]) - 2 + k[1])
    diff = [abs(int(rc[i]) - ro[k]) for i, k in enumerate(keys)]
    assert max(diff) <= 1 and sum(d > 0 for d in diff) <= 1, diff
    H.oracle_free(oc)
    lib.free_tree_hodlr(C.byref(ha))
    lib.free_tree_hodlr(C.byref(hb))
    for t in (ta, tb, tc):
        t.free()


def test_shape_mismatch_is_an_input_error(gpu):
    """Trees of different leaf sizes are refused by the device API."""
    from hmat_lib_amd import device as D
    n = 64
    pts = np.sort(WL.uniform(9, n))[:, None]
    da = D.fill_kernel_matrix(0, pts, 0.1, 0.0)
    t1 = D.DeviceTree.compress(da, n, 2, 1e-8)
    t2 = D.DeviceTree.compress(da, n, 2, 1e-8, ms=[10, 20, 14, 20])
    out = C.c_void_p()
    rc = capi.hip().hm_hip_multiply_hodlr_hodlr(t1.h, t2.h, 1e-8, C.byref(out), None)
    assert rc == capi.INPUT_ERROR and not out.value
    t1.free()
    t2.free()
