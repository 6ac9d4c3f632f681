"""ctypes view of the hmat_lib C API and of the hm_hip device C-ABI.

The struct classes mirror include/hmat_lib/hodlr.h (the reference layout,
include/hmat_lib/hodlr.h:7-263 of hmat_lib) so Python code can walk a tree
built by any library with that ABI: the MI355X product (libhmat_lib.so) or
a CPU checker.  Loading fails loudly when the built libraries are missing.
"""
from __future__ import annotations

import ctypes as C
import os
from ctypes import POINTER, c_double, c_int, c_long, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

SUCCESS, ALLOCATION_FAILURE, SVD_FAILURE, SVD_ALLOCATION_FAILURE, INPUT_ERROR, QR_FAILURE = range(6)
ERROR_NAMES = ["SUCCESS", "ALLOCATION_FAILURE", "SVD_FAILURE",
               "SVD_ALLOCATION_FAILURE", "INPUT_ERROR", "QR_FAILURE"]
DIAGONAL, OFFDIAGONAL, INTERNAL = range(3)


class NodeDiagonal(C.Structure):
    _fields_ = [("data", POINTER(c_double)), ("m", c_int)]


class NodeOffDiagonal(C.Structure):
    _fields_ = [("u", POINTER(c_double)), ("v", POINTER(c_double)),
                ("m", c_int), ("s", c_int), ("n", c_int)]


class HODLRData(C.Union):
    _fields_ = [("diagonal", NodeDiagonal), ("off_diagonal", NodeOffDiagonal)]


class HODLRLeafNode(C.Structure):
    pass


class HODLRInternalNode(C.Structure):
    pass


class HODLRNode(C.Union):
    _fields_ = [("leaf", POINTER(HODLRLeafNode)),
                ("internal", POINTER(HODLRInternalNode))]


HODLRLeafNode._fields_ = [("type", c_int), ("data", HODLRData),
                          ("parent", POINTER(HODLRInternalNode))]
HODLRInternalNode._fields_ = [("children", HODLRNode * 4),
                              ("parent", POINTER(HODLRInternalNode)),
                              ("m", c_int)]


class TreeHODLR(C.Structure):
    _fields_ = [("height", c_int),
                ("root", POINTER(HODLRInternalNode)),
                ("innermost_leaves", POINTER(POINTER(HODLRLeafNode))),
                ("len_work_queue", c_long),
                ("work_queue", POINTER(POINTER(HODLRInternalNode))),
                ("memory_internal_ptr", c_void_p),
                ("memory_leaf_ptr", c_void_p)]


PTree = POINTER(TreeHODLR)
PNode = POINTER(HODLRInternalNode)
PDouble = POINTER(c_double)
PInt = POINTER(c_int)


class TreeInfo(C.Structure):
    _fields_ = [("height", c_int), ("n", c_int), ("nleaf", c_int),
                ("nblock", c_int), ("max_rank", c_int),
                ("sum_rank_dims", C.c_longlong), ("diag_elems", C.c_longlong)]


class CompressOpts(C.Structure):
    _fields_ = [("sketch_init", c_int), ("sketch_max", c_int), ("n_test", c_int),
                ("seed", C.c_ulonglong), ("rank_hint", PInt),
                ("oversample", c_int), ("accept", c_double),
                ("rank_model", c_void_p), ("block_features", POINTER(C.c_float)),
                ("n_extra", c_int), ("margin", c_double), ("first_sigma", POINTER(c_double)),
                ("need_out", PInt)]


class CompressStats(C.Structure):
    _fields_ = [("sketch_rounds", c_int), ("blocks_retried", c_int),
                ("max_sketch", c_int), ("fallback_blocks", c_int),
                ("ms_total", c_double), ("predicted_blocks", c_int),
                ("predicted_failed", c_int), ("continued_blocks", c_int)]


def dptr(a: np.ndarray):
    """Pointer to the data of a C- or F-contiguous float64 array."""
    assert a.dtype == np.float64
    return a.ctypes.data_as(PDouble)


def _declare_hmat_api(lib, test_abi=False):
    """argtypes of the hmat_lib API (include/hmat_lib/*.h)."""
    mal = C.CFUNCTYPE(c_void_p, C.c_size_t)
    fre = C.CFUNCTYPE(None, c_void_p)
    lib.allocate_tree.argtypes = [c_int, PInt]
    lib.allocate_tree.restype = PTree
    if test_abi:
        lib.allocate_tree_monolithic.argtypes = [c_int, PInt, mal, fre]
        lib.free_tree_hodlr.argtypes = [POINTER(PTree), fre]
        lib.free_tree_data.argtypes = [PTree, fre]
        lib.dense_to_tree_hodlr.argtypes = [PTree, c_int, PInt, PDouble, c_double,
                                            PInt, mal, fre]
    else:
        lib.allocate_tree_monolithic.argtypes = [c_int, PInt]
        lib.free_tree_hodlr.argtypes = [POINTER(PTree)]
        lib.free_tree_data.argtypes = [PTree]
        lib.dense_to_tree_hodlr.argtypes = [PTree, c_int, PInt, PDouble, c_double, PInt]
    lib.allocate_tree_monolithic.restype = PTree
    lib.free_tree_hodlr.restype = None
    lib.free_tree_data.restype = None
    lib.dense_to_tree_hodlr.restype = c_int
    lib.compute_construct_tree_array_sizes.argtypes = [c_int] + [POINTER(C.c_size_t)] * 4
    lib.compute_construct_tree_array_sizes.restype = None
    lib.construct_tree.argtypes = [c_int, PTree, PNode, POINTER(HODLRLeafNode),
                                   POINTER(PNode), POINTER(POINTER(HODLRLeafNode)), PInt]
    lib.construct_tree.restype = None
    lib.multiply_vector.argtypes = [PTree, PDouble, PDouble, PInt]
    lib.multiply_vector.restype = PDouble
    for name in ("multiply_hodlr_dense", "multiply_hodlr_transpose_dense",
                 "multiply_dense_hodlr"):
        f = getattr(lib, name)
        f.argtypes = [PTree, PDouble, c_int, c_int, PDouble, c_int, PInt]
        f.restype = PDouble
    for name in ("multiply_internal_node_dense", "multiply_internal_node_transpose_dense",
                 "multiply_dense_internal_node"):
        f = getattr(lib, name)
        f.argtypes = [PNode, c_int, PDouble, c_int, c_int, POINTER(PNode), PDouble,
                      PDouble, c_int]
        f.restype = None
    lib.compute_multiply_hodlr_dense_workspace.argtypes = [PTree, c_int]
    lib.compute_multiply_hodlr_dense_workspace.restype = c_int
    lib.get_highest_s.argtypes = [PTree]
    lib.get_highest_s.restype = c_int
    lib.multiply_hodlr_hodlr.argtypes = [PTree, PTree, PTree, c_double, PInt]
    lib.multiply_hodlr_hodlr.restype = c_int


def _declare_hip_api(lib):
    V = c_void_p
    sig = {
        "hm_hip_init": ([c_int], c_int),
        "hm_hip_device_count": ([PInt], c_int),
        "hm_hip_last_error": ([], C.c_char_p),
        "hm_hip_synchronize": ([V], c_int),
        "hm_hip_release_buffers": ([], c_int),
        "hm_hip_malloc": ([POINTER(V), C.c_size_t], c_int),
        "hm_hip_free": ([V], c_int),
        "hm_hip_memcpy_h2d": ([V, V, C.c_size_t, V], c_int),
        "hm_hip_memcpy_d2h": ([V, V, C.c_size_t, V], c_int),
        "hm_hip_memcpy_d2d": ([V, V, C.c_size_t, V], c_int),
        "hm_hip_copy2d_h2d": ([V, c_long, V, c_long, c_int, c_int, V], c_int),
        "hm_hip_copy2d_d2h": ([V, c_long, V, c_long, c_int, c_int, V], c_int),
        "hm_hip_tree_upload": ([PNode, c_int, POINTER(V), V], c_int),
        "hm_hip_tree_download": ([V, PTree, V, V], c_int),
        "hm_hip_tree_destroy": ([V], c_int),
        "hm_hip_tree_info": ([V, POINTER(TreeInfo)], c_int),
        "hm_hip_tree_ranks": ([V, PInt], c_int),
        "hm_hip_matvec": ([V, V, V, V], c_int),
        "hm_hip_matmat": ([V, c_int, c_int, V, c_long, c_long, V, c_long, c_long, V], c_int),
        "hm_hip_matmat_noalloc": ([V, c_int, c_int, V, c_long, c_long, V, c_long, c_long, V],
                                  c_int),
        "hm_hip_compress_opts_default": ([POINTER(CompressOpts)], None),
        "hm_hip_compress_dense": ([V, c_long, c_int, c_int, PInt, c_double,
                                   POINTER(CompressOpts), POINTER(V),
                                   POINTER(CompressStats), V], c_int),
        "hm_hip_fill_kernel_matrix": ([c_int, c_int, c_int, V, c_double, c_double,
                                       V, c_long, V], c_int),
        "hm_hip_recompress_batch": ([c_int, PInt, PInt, PInt, POINTER(V), POINTER(V),
                                     c_double, c_int, PInt, V], c_int),
        "hm_hip_multiply_hodlr_hodlr": ([V, V, c_double, POINTER(V), V], c_int),
        "hm_hip_tree_save": ([V, C.c_char_p, V], c_int),
        "hm_hip_tree_load": ([C.c_char_p, POINTER(V), V], c_int),
        "hm_hip_selftest_mfma": ([PDouble], c_int),
        "hm_hip_selftest_mfma_bf16": ([POINTER(C.c_float), POINTER(C.c_float),
                                       POINTER(C.c_float)], c_int),
        "hm_hip_selftest_mfma_lowp": ([c_int, POINTER(C.c_float), POINTER(C.c_float),
                                       POINTER(C.c_float)], c_int),
        "hm_hip_mlp_create": ([c_int, c_int, C.c_ulonglong, POINTER(V)], c_int),
        "hm_hip_mlp_destroy": ([V], c_int),
        "hm_hip_mlp_nparams": ([V], C.c_longlong),
        "hm_hip_mlp_get_params": ([V, POINTER(C.c_float)], c_int),
        "hm_hip_mlp_set_params": ([V, POINTER(C.c_float)], c_int),
        "hm_hip_mlp_forward": ([V, V, c_int, V, V], c_int),
        "hm_hip_mlp_gradient": ([V, V, V, c_int, POINTER(C.c_float), POINTER(C.c_float), V],
                                c_int),
        "hm_hip_mlp_train": ([V, V, V, c_int, c_int, c_int, C.c_float, C.c_ulonglong,
                              POINTER(C.c_float), V], c_int),
        "hm_hip_rank_features": ([c_int, c_int, POINTER(C.c_float), POINTER(c_double), c_double,
                                  POINTER(C.c_float)], c_int),
        "hm_hip_compress_batch": ([c_int, POINTER(V), POINTER(c_long), POINTER(c_int),
                                   POINTER(c_int), c_double, V, POINTER(V), V, V], c_int),
        "hm_hip_batch_create": ([c_int, POINTER(V), c_int, POINTER(c_int), POINTER(V),
                                 POINTER(c_long), POINTER(V), POINTER(c_long), POINTER(V)],
                                c_int),
        "hm_hip_batch_tune": ([V, V], c_int),
        "hm_hip_batch_run": ([V, c_int, V], c_int),
        "hm_hip_batch_variants": ([V, POINTER(c_int), c_int], c_int),
        "hm_hip_batch_destroy": ([V], c_int),
        "hm_hip_tree_set_precision": ([V, c_int, V], c_int),
        "hm_hip_tree_set_format": ([V, c_int, c_int, V], c_int),
        "hm_hip_kernel_timing_dump": ([C.c_char_p, c_int], c_int),
        "hm_hip_qr_batch": ([c_int, PInt, PInt, POINTER(V), POINTER(c_long), POINTER(V),
                             POINTER(c_long), POINTER(V), POINTER(c_long), V], c_int),
        "hm_hip_audit_qr": ([c_int, POINTER(c_int), POINTER(c_int), c_int, C.c_char_p,
                             c_int], c_int),
        "hm_hip_audit_hh": ([c_int, POINTER(c_int), POINTER(c_int), POINTER(c_int), C.c_char_p,
                             c_int], c_int),
        "hm_hip_matmat_variant": ([V, c_int, c_int], c_int),
        "hm_hip_kernel_timing": ([c_int], c_int),
        "hm_hip_kernel_timing_select": ([C.c_char_p], c_int),
        "hm_hip_kernel_timing_read": ([C.c_char_p, PDouble, POINTER(c_long)], c_int),
        "hm_hip_kernel_timing_reset": ([], c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res


_LIBS = {}


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise RuntimeError(
            f"hmat_lib_amd: {path} is not built; run __graft_entry__.build() "
            "(make -C hmat_lib_amd/csrc) first")
    return C.CDLL(path, mode=C.RTLD_LOCAL)


def hip():
    """libhmat_hip.so (hm_hip_* device C-ABI)."""
    if "hip" not in _LIBS:
        lib = _load("libhmat_hip.so")
        _declare_hip_api(lib)
        _LIBS["hip"] = lib
    return _LIBS["hip"]


def hmat(test_abi=False):
    """libhmat_lib.so: the hmat_lib drop-in API (or its _TEST_HODLR build)."""
    key = "hmat_test" if test_abi else "hmat"
    if key not in _LIBS:
        hip()
        lib = _load("libhmat_lib_test.so" if test_abi else "libhmat_lib.so")
        _declare_hmat_api(lib, test_abi=test_abi)
        lib.hmat_amd_invalidate.argtypes = [PTree]
        lib.hmat_amd_invalidate.restype = None
        lib.hmat_amd_device_tree.argtypes = [PTree, PInt]
        lib.hmat_amd_device_tree.restype = c_void_p
        lib.hmat_amd_attach.argtypes = [PTree, c_void_p, PInt]
        lib.hmat_amd_attach.restype = None
        lib.hmat_amd_cache_size.argtypes = []
        lib.hmat_amd_cache_size.restype = c_int
        lib.hmat_amd_save_tree.argtypes = [PTree, C.c_char_p]
        lib.hmat_amd_save_tree.restype = c_int
        lib.hmat_amd_load_tree.argtypes = [C.c_char_p, PInt]
        lib.hmat_amd_load_tree.restype = PTree
        lib.hmat_amd_read_dense_text.argtypes = [C.c_char_p, PInt, PInt]
        lib.hmat_amd_read_dense_text.restype = PDouble
        recomp = [C.POINTER(NodeOffDiagonal), c_int, c_double, PInt]
        lib.recompress.argtypes = recomp
        lib.recompress.restype = c_int
        lib.recompress_large_s.argtypes = recomp
        lib.recompress_large_s.restype = c_int
        _LIBS[key] = lib
    return _LIBS[key]


def last_error() -> str:
    return hip().hm_hip_last_error().decode(errors="replace")


def check(rc: int, what: str = "hm_hip call"):
    if rc != SUCCESS:
        raise RuntimeError(f"{what} failed ({ERROR_NAMES[rc] if 0 <= rc < 6 else rc}): "
                           f"{last_error()}")


def declare_hmat_api(lib, test_abi=False):
    """Set argtypes of the hmat_lib API on any CDLL with that ABI."""
    _declare_hmat_api(lib, test_abi=test_abi)
    return lib
