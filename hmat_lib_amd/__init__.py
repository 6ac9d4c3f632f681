"""hmat_lib_amd: MI355X-native HODLR hot path behind hmat_lib's C API.

libhmat_lib.so   the hmat_lib API (allocators, dense_to_tree_hodlr,
                 multiply_vector, multiply_hodlr_dense and variants,
                 recompression), host C over
libhmat_hip.so   the hm_hip C-ABI: hand-written gfx950 kernels.

Python modules: capi (ctypes bindings), tree (walk/extract trees),
device (device-resident path), workloads (seeded synthetic inputs),
build (compile the libraries).
"""
from . import capi  # noqa: F401

__all__ = ["capi", "tree", "device", "workloads", "build"]
