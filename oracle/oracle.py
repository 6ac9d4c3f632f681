"""ORACLE - TEST INFRASTRUCTURE ONLY (checker, never the thing measured).

ctypes bindings of
  oracle/lib/liboracle.so   the CPU restatement (hodlr_oracle.c), and
  oracle/_ref/libhmat_ref.so the reference hmat_lib built from its own
                             sources (present only where it was built).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline use this.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

os.environ.setdefault("MKL_THREADING_LAYER", "GNU")

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from hmat_lib_amd import capi  # noqa: E402
from hmat_lib_amd.capi import PDouble, PInt, PNode, PTree, c_double, c_int  # noqa: E402

ORACLE_SO = os.path.join(HERE, "lib", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhmat_ref.so")

_cache = {}


def oracle():
    if "o" not in _cache:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
        lib = C.CDLL(ORACLE_SO, mode=C.RTLD_LOCAL)
        lib.orc_allocate_tree.argtypes = [c_int, PInt]
        lib.orc_allocate_tree.restype = PTree
        lib.orc_free_tree.argtypes = [PTree]
        lib.orc_free_tree.restype = None
        lib.orc_free_tree_data.argtypes = [PTree]
        lib.orc_free_tree_data.restype = None
        lib.orc_block_sizes_halves.argtypes = [PTree, c_int]
        lib.orc_block_sizes_custom.argtypes = [PTree, PInt]
        lib.orc_dense_to_tree.argtypes = [PTree, c_int, PInt, PDouble, c_double, PInt]
        lib.orc_dense_to_tree.restype = c_int
        lib.orc_compress_block.argtypes = [c_int, c_int, PDouble, C.c_long, c_double,
                                           C.POINTER(PDouble), C.POINTER(PDouble), PInt,
                                           PDouble, PInt]
        lib.orc_compress_block.restype = c_int
        lib.orc_multiply_vector.argtypes = [PTree, PDouble, PDouble, PInt]
        lib.orc_multiply_vector.restype = PDouble
        for name in ("orc_multiply_hodlr_dense", "orc_multiply_hodlr_transpose_dense",
                     "orc_multiply_dense_hodlr"):
            f = getattr(lib, name)
            f.argtypes = [PTree, PDouble, c_int, c_int, PDouble, c_int, PInt]
            f.restype = PDouble
        lib.orc_multiply_internal_node_dense.argtypes = [PNode, c_int, c_int, PDouble,
                                                         c_int, c_int, PDouble, c_int]
        lib.orc_multiply_internal_node_dense.restype = None
        lib.orc_multiply_dense_internal_node.argtypes = [PNode, c_int, PDouble, c_int,
                                                         c_int, PDouble, c_int]
        lib.orc_multiply_dense_internal_node.restype = None
        lib.orc_get_highest_s.argtypes = [PTree]
        lib.orc_get_highest_s.restype = c_int
        lib.orc_tree_to_dense.argtypes = [PTree, PDouble, C.c_long]
        lib.orc_tree_to_dense.restype = None
        rec = [C.POINTER(capi.NodeOffDiagonal), c_int, c_double, PInt]
        lib.orc_recompress.argtypes = rec
        lib.orc_recompress.restype = c_int
        lib.orc_recompress_large_s.argtypes = rec
        lib.orc_recompress_large_s.restype = c_int
        lib.orc_multiply_hodlr_hodlr.argtypes = [PTree, PTree, PTree, c_double, PInt]
        lib.orc_multiply_hodlr_hodlr.restype = c_int
        lib.orc_fill_kernel_block.argtypes = [c_int, c_int, PDouble, c_double, c_double,
                                              C.c_long, C.c_long, C.c_long, C.c_long,
                                              PDouble, C.c_long]
        lib.orc_fill_kernel_block.restype = None
        _cache["o"] = lib
    return _cache["o"]


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def ref():
    """The reference hmat_lib (unmodified sources) as a checker library."""
    if "r" not in _cache:
        if not have_ref():
            raise RuntimeError(f"{REF_SO} missing (reference sources absent at build time)")
        lib = C.CDLL(REF_SO, mode=C.RTLD_LOCAL)
        capi.declare_hmat_api(lib)
        lib.ref_compress_off_diagonal.argtypes = [C.POINTER(capi.NodeOffDiagonal), c_int,
                                                  c_int, PDouble, PDouble, PDouble,
                                                  PDouble, c_double, PInt]
        lib.ref_compress_off_diagonal.restype = c_int
        rec = [C.POINTER(capi.NodeOffDiagonal), c_int, c_double, PInt]
        lib.ref_recompress.argtypes = rec
        lib.ref_recompress.restype = c_int
        lib.ref_recompress_large_s.argtypes = rec
        lib.ref_recompress_large_s.restype = c_int
        _cache["r"] = lib
    return _cache["r"]


def libc():
    if "c" not in _cache:
        lib = C.CDLL("libc.so.6")
        lib.srand.argtypes = [C.c_uint]
        lib.rand.restype = c_int
        _cache["c"] = lib
    return _cache["c"]
