"""Host-side logic of the drop-in library (no GPU compute): symbol exports,
tree allocators, size formulas, rank/workspace queries, input errors.
CPU only."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from tests import helpers as H

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(header).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", txt))
    return {n for n in names if not n.startswith(("if", "defined", "sizeof", "__attribute__",
                                                  "visibility", "void", "int", "double",
                                                  "struct", "return"))}


def _exports(path):
    out = os.popen(f"nm -D --defined-only {path}").read()
    return {l.split()[-1] for l in out.splitlines() if " T " in l}


def test_every_declared_symbol_is_exported():
    hip_syms = _exports(os.path.join(capi.LIBDIR, "libhmat_hip.so"))
    lib_syms = _exports(os.path.join(capi.LIBDIR, "libhmat_lib.so"))
    test_syms = _exports(os.path.join(capi.LIBDIR, "libhmat_lib_test.so"))
    for h in glob.glob(os.path.join(ROOT, "include", "hmat_lib", "*.h")) + \
            [os.path.join(ROOT, "include", "hmat_lib_amd.h")]:
        for name in _declared(h):
            if name in ("malloc", "free", "alloc", "permanent_allocator"):
                continue
            assert name in lib_syms, (h, name)
            assert name in test_syms, (h, name)
    for name in _declared(os.path.join(ROOT, "include", "hm_hip.h")):
        if name.startswith("hm_hip_"):
            assert name in hip_syms, name


def test_libraries_load():
    capi.hip()
    capi.hmat()
    capi.hmat(test_abi=True)


def _consistent(t, height):
    tc = t.contents
    assert tc.height == height and tc.len_work_queue == 1 << (height - 1)
    assert not tc.root.contents.parent
    lev = T.levels(tc.root, height)
    for d in range(height):
        for p in lev[d]:
            for c in (1, 2):
                leaf = p.contents.children[c].leaf.contents
                assert leaf.type == capi.OFFDIAGONAL
                assert C.addressof(leaf.parent.contents) == C.addressof(p.contents)
                od = leaf.data.off_diagonal
                assert not od.u and not od.v and od.m == od.s == od.n == 0
            if d + 1 < height:
                for c in (0, 3):
                    ch = p.contents.children[c].internal.contents
                    assert C.addressof(ch.parent.contents) == C.addressof(p.contents)
                    assert ch.m == 0
    for q, p in enumerate(lev[-1]):
        assert C.addressof(tc.work_queue[q].contents) == C.addressof(p.contents)
        for k, c in enumerate((0, 3)):
            leaf = p.contents.children[c].leaf
            assert leaf.contents.type == capi.DIAGONAL
            assert not leaf.contents.data.diagonal.data
            assert C.addressof(tc.innermost_leaves[2 * q + k].contents) == \
                C.addressof(leaf.contents)


@pytest.mark.parametrize("height", range(1, 7))
def test_allocators_structure(height):
    lib = capi.hmat()
    ierr = C.c_int(-1)
    for fn in (lib.allocate_tree, lib.allocate_tree_monolithic):
        t = fn(height, C.byref(ierr))
        assert ierr.value == capi.SUCCESS
        _consistent(t, height)
        lib.free_tree_hodlr(C.byref(t))
        assert not t


def test_array_sizes():
    """test_allocators.c:82-99."""
    lib = capi.hmat()
    exp = [(1, 1, 2, 4), (2, 3, 4, 10), (4, 7, 8, 22), (8, 15, 16, 46), (16, 31, 32, 94),
           (32, 63, 64, 190)]
    for h, (wq, ni, il, nl) in enumerate(exp, start=1):
        s = [C.c_size_t() for _ in range(4)]
        lib.compute_construct_tree_array_sizes(h, *[C.byref(x) for x in s])
        si, sl, sw, sil = (x.value for x in s)
        assert sw == wq * 8 and sil == il * 8
        assert si == ni * C.sizeof(capi.HODLRInternalNode)
        assert sl == nl * C.sizeof(capi.HODLRLeafNode)


def test_struct_layout_matches_reference_abi():
    # sizes on LP64 of the reference structs (pointer 8, int 4, padding)
    assert C.sizeof(capi.NodeOffDiagonal) == 32
    assert C.sizeof(capi.HODLRLeafNode) == 48
    assert C.sizeof(capi.HODLRInternalNode) == 48
    assert C.sizeof(capi.TreeHODLR) == 56


def test_construct_tree_from_arrays():
    lib = capi.hmat()
    h = 3
    s = [C.c_size_t() for _ in range(4)]
    lib.compute_construct_tree_array_sizes(h, *[C.byref(x) for x in s])
    tree = capi.TreeHODLR()
    internal = (C.c_char * s[0].value)()
    leaves = (C.c_char * s[1].value)()
    wq = (C.c_char * s[2].value)()
    il = (C.c_char * s[3].value)()
    ierr = C.c_int(-1)
    lib.construct_tree(h, C.byref(tree), C.cast(internal, capi.PNode),
                       C.cast(leaves, C.POINTER(capi.HODLRLeafNode)),
                       C.cast(wq, C.POINTER(capi.PNode)),
                       C.cast(il, C.POINTER(C.POINTER(capi.HODLRLeafNode))), C.byref(ierr))
    assert ierr.value == capi.SUCCESS
    _consistent(C.pointer(tree), h)
    assert C.addressof(tree.root.contents) == C.addressof(internal)


def test_input_errors():
    lib = capi.hmat()
    ierr = C.c_int(-1)
    assert not lib.allocate_tree(0, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    assert not lib.allocate_tree_monolithic(0, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    t = lib.allocate_tree_monolithic(2, C.byref(ierr))
    a = np.zeros((8, 8), order="F")
    lib.dense_to_tree_hodlr(None, 8, None, capi.dptr(a), 1e-8, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    lib.dense_to_tree_hodlr(t, 8, None, None, 1e-8, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    ms = (C.c_int * 4)(2, 2, 2, 1)  # sums to 7, not 8
    lib.dense_to_tree_hodlr(t, 8, ms, capi.dptr(a), 1e-8, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    x = np.zeros(8)
    assert not lib.multiply_vector(t, capi.dptr(x), capi.dptr(x), C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    assert not lib.multiply_hodlr_dense(t, None, 1, 8, None, 8, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    lib.free_tree_hodlr(C.byref(t))


def test_highest_rank_and_workspace():
    """test_hodlr_dense_algebra.c:51-113: workspace = max rank x matrix_a."""
    lib = capi.hmat()
    rng = np.random.default_rng(0)
    for h in range(1, 5):
        ft = H.random_flat_tree(rng, [8] * (1 << h), lambda l, j: h)
        keep = T.Keep()
        t = T.fill_tree(lib, ft, keep)
        assert lib.get_highest_s(t) == h
        assert lib.compute_multiply_hodlr_dense_workspace(t, 2048) == h * 2048
        T.detach_and_free(lib, t)
        ft = H.random_flat_tree(rng, [8] * (1 << h),
                                lambda l, j, h=h: (4 + h) if (l == h and j == 1) else 1)
        t = T.fill_tree(lib, ft, keep)
        assert lib.compute_multiply_hodlr_dense_workspace(t, 42) == (4 + h) * 42
        T.detach_and_free(lib, t)


def test_test_abi_uses_injected_allocator():
    lib = capi.hmat(test_abi=True)
    libc = C.CDLL("libc.so.6")
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    libc.free.argtypes = [C.c_void_p]
    counts = {"m": 0, "f": 0}
    MAL = C.CFUNCTYPE(C.c_void_p, C.c_size_t)
    FRE = C.CFUNCTYPE(None, C.c_void_p)

    @MAL
    def mal(n):
        counts["m"] += 1
        return libc.malloc(n)

    @FRE
    def fre(p):
        counts["f"] += 1
        libc.free(p)

    ierr = C.c_int(-1)
    t = lib.allocate_tree_monolithic(3, C.byref(ierr), mal, fre)
    assert ierr.value == 0 and counts["m"] == 5
    _consistent(t, 3)
    lib.free_tree_hodlr(C.byref(t), fre)
    assert counts["f"] == 5


@pytest.mark.skipif(os.environ.get("HM_FORCE_GPU_ABSENT_TEST") is None and
                    __import__("hmat_lib_amd.device", fromlist=["x"]).device_count() > 0,
                    reason="a GPU is present")
def test_no_device_fails_loudly(capfd):
    """Without a HIP device the compute path reports ALLOCATION_FAILURE with
    a message; there is no host fallback."""
    lib = capi.hmat()
    ierr = C.c_int(-1)
    t = lib.allocate_tree_monolithic(1, C.byref(ierr))
    a = np.asfortranarray(np.eye(4))
    lib.dense_to_tree_hodlr(t, 4, None, capi.dptr(a), 1e-8, C.byref(ierr))
    assert ierr.value == capi.ALLOCATION_FAILURE
    err = capfd.readouterr().err
    assert "hmat_lib (MI355X)" in err and "failed" in err
    lib.free_tree_hodlr(C.byref(t))


QR_SHAPES = [
    # (m, L): single panel, exact multiples of the 64-column panel, ragged
    # last panel, square, long K that triggers split-K projections
    [(1, 1), (7, 7), (64, 64), (65, 64)],
    [(200, 65), (300, 128), (129, 129), (1000, 200)],
    [(32768, 518), (16384, 390), (5000, 64), (70, 1)],
]


@pytest.mark.parametrize("with_r", [0, 1])
@pytest.mark.parametrize("shapes", QR_SHAPES, ids=["small", "ragged", "wide"])
def test_qr_program_stays_in_bounds(shapes, with_r):
    """Every work item of the wide-panel QR launch program (projections,
    TSQR levels, copies, triangular-factor updates) reads and writes only
    inside the buffers it was given; checked on the host, no device."""
    import ctypes as C
    from hmat_lib_amd import capi
    m = (C.c_int * len(shapes))(*[s[0] for s in shapes])
    L = (C.c_int * len(shapes))(*[s[1] for s in shapes])
    msg = C.create_string_buffer(256)
    rc = capi.hip().hm_hip_audit_qr(len(shapes), m, L, with_r, msg, 256)
    assert rc == 0, msg.value.decode()


HH_CASES = [
    (1, [10, 11], None),
    (3, [7, 12, 5, 9, 13, 8, 6, 10], None),
    (4, [6, 0, 9, 1, 14, 3, 8, 8, 2, 11, 5, 7, 13, 4, 9, 1], None),  # empty and 1-row leaves
    (5, [64] * 32, "wide"),  # ranks up to the leaf size: dense recompression
]


@pytest.mark.parametrize("height,leaves,kind", HH_CASES, ids=["h1", "h3", "ragged", "wide"])
def test_hh_program_stays_in_bounds(height, leaves, kind):
    """Every work item of the HODLR x HODLR launch program (T cores,
    subtree W pieces, leaf and ancestor GEMMs, chained GEMMs, copies) reads
    and writes only inside the two trees, the product's diagonal, the
    pre-recompression factor pools and its scratch; checked on the host."""
    import ctypes as C
    import numpy as np
    from hmat_lib_amd import capi
    rng = np.random.default_rng(height)
    nb = (1 << (height + 1)) - 2
    hi = 64 if kind == "wide" else 6
    lv = (C.c_int * len(leaves))(*leaves)
    ra = (C.c_int * nb)(*[int(x) for x in rng.integers(1, hi, size=nb)])
    rb = (C.c_int * nb)(*[int(x) for x in rng.integers(1, hi, size=nb)])
    msg = C.create_string_buffer(256)
    rc = capi.hip().hm_hip_audit_hh(height, lv, ra, rb, msg, 256)
    assert rc == 0, msg.value.decode()


def test_hh_dropin_argument_errors():
    """multiply_hodlr_hodlr: NULL operands and trees of different heights
    give INPUT_ERROR before any device work."""
    import ctypes as C
    from hmat_lib_amd import capi
    lib = capi.hmat()
    ierr = C.c_int(0)
    t2 = lib.allocate_tree(2, C.byref(ierr))
    t3 = lib.allocate_tree(3, C.byref(ierr))
    assert lib.multiply_hodlr_hodlr(t2, t3, t2, 1e-8, C.byref(ierr)) == 0
    assert ierr.value == capi.INPUT_ERROR
    ierr.value = 0
    lib.multiply_hodlr_hodlr(None, t2, t2, 1e-8, C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR
    lib.free_tree_hodlr(C.byref(t2))
    lib.free_tree_hodlr(C.byref(t3))


def _flat_equal(fa, fb):
    import numpy as np
    assert fa.height == fb.height and list(fa.leaf_sizes) == list(fb.leaf_sizes)
    for d1, d2 in zip(fa.diag, fb.diag):
        np.testing.assert_array_equal(d1, d2)
    assert fa.blocks.keys() == fb.blocks.keys()
    for k in fa.blocks:
        np.testing.assert_array_equal(fa.blocks[k][0], fb.blocks[k][0])
        np.testing.assert_array_equal(fa.blocks[k][1], fb.blocks[k][1])


def test_binary_tree_round_trip(tmp_path):
    """save -> load gives back every array bit for bit (uneven and empty
    leaves, ranks 0..4), the internal node sizes, and a file of exactly
    header + 8 * (sum m_c^2 + sum (m_b + n_b) s_b) bytes."""
    import ctypes as C
    import numpy as np
    from hmat_lib_amd import capi
    from hmat_lib_amd import tree as T
    from tests import helpers as H
    lib = capi.hmat()
    rng = np.random.default_rng(4)
    leaves = [7, 0, 12, 5, 9, 1, 6, 10]
    ft = H.random_flat_tree(rng, leaves, lambda l, j: int(rng.integers(0, 5)))
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    path = str(tmp_path / "t.hodlr")
    assert lib.hmat_amd_save_tree(t, path.encode()) == capi.SUCCESS
    nblk = sum(1 << l for l in range(1, 4))
    head = (32 + 4 * (8 + nblk) + 7) // 8 * 8
    payload = sum(m * m for m in leaves)
    for (l, j), (u, v) in ft.blocks.items():
        payload += (u.shape[0] + v.shape[0]) * u.shape[1]
    assert (tmp_path / "t.hodlr").stat().st_size == head + 8 * payload
    ierr = C.c_int(-1)
    t2 = lib.hmat_amd_load_tree(path.encode(), C.byref(ierr))
    assert ierr.value == capi.SUCCESS and t2
    _flat_equal(ft, T.flatten(t2))
    lev = T.levels(t2.contents.root, 3)
    assert t2.contents.root.contents.m == sum(leaves)
    assert [p.contents.m for p in lev[1]] == [sum(leaves[:4]), sum(leaves[4:])]
    lib.free_tree_hodlr(C.byref(t2))
    T.detach_and_free(lib, t)


def test_binary_tree_rejects_bad_files(tmp_path):
    """Wrong magic, a truncated payload, leaf sizes that do not add up and a
    missing file are INPUT_ERROR with no tree."""
    import ctypes as C
    import struct
    import numpy as np
    from hmat_lib_amd import capi
    from hmat_lib_amd import tree as T
    from tests import helpers as H
    lib = capi.hmat()
    rng = np.random.default_rng(5)
    ft = H.random_flat_tree(rng, [4, 3, 5, 2], lambda l, j: 1)
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    good = tmp_path / "good.hodlr"
    assert lib.hmat_amd_save_tree(t, str(good).encode()) == capi.SUCCESS
    T.detach_and_free(lib, t)
    data = good.read_bytes()
    cases = {"magic": b"XODLRBIN" + data[8:], "truncated": data[:-8],
             "longer": data + b"\0" * 8,
             "sum": data[:24] + struct.pack("<q", 15) + data[32:]}
    for name, blob in cases.items():
        p = tmp_path / f"{name}.hodlr"
        p.write_bytes(blob)
        ierr = C.c_int(0)
        assert not lib.hmat_amd_load_tree(str(p).encode(), C.byref(ierr)), name
        assert ierr.value == capi.INPUT_ERROR, name
    ierr = C.c_int(0)
    assert not lib.hmat_amd_load_tree(str(tmp_path / "none").encode(), C.byref(ierr))
    assert ierr.value == capi.INPUT_ERROR


def test_read_dense_text_reference_format(tmp_path):
    """The reference's test format (tests/utils/io.c:15-81): one row per
    line, spaces or tabs; read into column-major storage.  A file without a
    newline or with missing entries is INPUT_ERROR."""
    import ctypes as C
    import numpy as np
    from hmat_lib_amd import capi
    lib = capi.hmat()
    a = np.random.default_rng(6).standard_normal((5, 5))
    lines = ["\t".join(f"{x:.17g}" for x in row) if i % 2 else
             "  ".join(f"{x:.17g}" for x in row) for i, row in enumerate(a)]
    p = tmp_path / "m.txt"
    p.write_text("\n".join(lines) + "\n")
    n, ierr = C.c_int(0), C.c_int(-1)
    ptr = lib.hmat_amd_read_dense_text(str(p).encode(), C.byref(n), C.byref(ierr))
    assert ierr.value == capi.SUCCESS and n.value == 5
    got = np.ctypeslib.as_array(ptr, shape=(25,)).reshape(5, 5, order="F").copy()
    np.testing.assert_array_equal(got, a)
    C.CDLL("libc.so.6").free(C.cast(ptr, C.c_void_p))
    for name, text in (("nonl", "1 2 3"), ("short", "1 2\n3\n")):
        q = tmp_path / name
        q.write_text(text)
        assert not lib.hmat_amd_read_dense_text(str(q).encode(), C.byref(n), C.byref(ierr))
        assert ierr.value == capi.INPUT_ERROR, name


def _block_features_loop(pts, height, ell):
    """The geometric inputs block by block, from their definition."""
    from hmat_lib_amd import learned as L
    pts = np.asarray(pts, dtype=np.float64)
    if pts.ndim == 1:
        pts = pts[:, None]
    off = L.leaf_offsets(pts.shape[0], height)
    rows = []
    for lvl in range(1, height + 1):
        span = 1 << (height - lvl)
        for j in range(1 << lvl):
            pr = pts[off[j * span]:off[(j + 1) * span]]
            pc = pts[off[(j ^ 1) * span]:off[((j ^ 1) + 1) * span]]
            cr = pr.mean(0) if len(pr) else np.zeros(pts.shape[1])
            cc = pc.mean(0) if len(pc) else np.zeros(pts.shape[1])
            dr = np.linalg.norm(pr.max(0) - pr.min(0)) if len(pr) else 0.0
            dc = np.linalg.norm(pc.max(0) - pc.min(0)) if len(pc) else 0.0
            rows.append([np.log2(max(len(pr), 1)) / 16, np.log2(max(len(pc), 1)) / 16, lvl / 16,
                         np.log10(np.linalg.norm(cr - cc) + 1e-3) / 3, np.log10(ell) / 2,
                         np.log10(dr + 1e-6) / 3, np.log10(dc + 1e-6) / 3])
    return np.asarray(rows, dtype=np.float32)


def test_block_features():
    """The learned predictor's geometric inputs: a known answer (8 points
    0..7 on a line, height 1: centroids 1.5 and 5.5, extents 3) and the
    per-block definition on 1-D and 2-D points, with empty leaves."""
    from hmat_lib_amd import learned as L
    from hmat_lib_amd import workloads as WL
    f = L.block_features(np.arange(8.0), 1, 0.1)
    want = [np.log2(4) / 16, np.log2(4) / 16, 1 / 16, np.log10(4 + 1e-3) / 3, np.log10(0.1) / 2,
            np.log10(3 + 1e-6) / 3, np.log10(3 + 1e-6) / 3]
    np.testing.assert_allclose(f, np.array([want, want], dtype=np.float32), rtol=1e-6)
    cases = [(np.sort(np.random.default_rng(1).uniform(size=300)), 5, 0.1),
             (WL.points_2d_morton(4, 1000), 3, 0.3), (WL.points_2d_morton(5, 20), 5, 0.3),
             (WL.points_2d_morton(6, 4096), 5, 0.05)]
    for pts, h, ell in cases:
        np.testing.assert_array_equal(L.block_features(pts, h, ell),
                                      _block_features_loop(pts, h, ell))
