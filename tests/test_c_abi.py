"""The drop-in seen from C: a C program compiled by gcc against
include/hmat_lib/*.h and linked to libhmat_lib.so (tests/c_abi/
conformance.c), and the public struct layout compiled from this
repository's headers vs the reference's own headers.

CPU: the host half of the C caller (tree skeletons, sizes, rank and
workspace queries) and the layout comparison (skipped where the
reference's headers are absent).  GPU: the full C caller (compression, a
walk of the tree through its public fields, all seven product entry
points, H x H), checked inside the program against exact dense algebra.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "hmat_lib_amd", "lib")
SRC = os.path.join(ROOT, "tests", "c_abi", "conformance.c")
REF_INC = "/root/reference/include"

# every field of every public struct / union of include/hmat_lib/hodlr.h
LAYOUT = {
    "struct NodeDiagonal": ["data", "m"],
    "struct NodeOffDiagonal": ["u", "v", "m", "s", "n"],
    "union HODLRData": ["diagonal", "off_diagonal"],
    "struct HODLRLeafNode": ["type", "data", "parent"],
    "union HODLRNode": ["leaf", "internal"],
    "struct HODLRInternalNode": ["children", "parent", "m"],
    "struct TreeHODLR": ["height", "root", "innermost_leaves", "len_work_queue",
                         "work_queue", "memory_internal_ptr", "memory_leaf_ptr"],
}
ENUMS = ["DIAGONAL", "OFFDIAGONAL", "INTERNAL", "SUCCESS", "ALLOCATION_FAILURE",
         "SVD_FAILURE", "SVD_ALLOCATION_FAILURE", "INPUT_ERROR", "QR_FAILURE"]


def _compile_caller(tmp_path):
    exe = os.path.join(str(tmp_path), "conformance")
    cmd = ["gcc", "-std=c11", "-O1", "-Wall", "-Werror", f"-I{INC}", SRC, f"-L{LIBDIR}",
           "-lhmat_lib", f"-Wl,-rpath,{LIBDIR}", "-lm", "-o", exe]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


def _layout_program(inc, tmp_path, tag):
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "hmat_lib/hodlr.h"',
             '#include "hmat_lib/error.h"', "int main(void) {"]
    for typ, fields in LAYOUT.items():
        lines.append(f'  printf("sizeof {typ} %zu\\n", sizeof({typ}));')
        for f in fields:
            lines.append(f'  printf("offsetof {typ}.{f} %zu\\n", offsetof({typ}, {f}));')
    for e in ENUMS:
        lines.append(f'  printf("enum {e} %d\\n", (int){e});')
    lines += ["  return 0;", "}"]
    src = os.path.join(str(tmp_path), f"layout_{tag}.c")
    with open(src, "w") as f:
        f.write("\n".join(lines) + "\n")
    exe = src[:-2]
    res = subprocess.run(["gcc", "-std=c2x", "-w", f"-I{inc}", src, "-o", exe],
                         capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    return out.splitlines()


def test_c_caller_host(tmp_path):
    exe = _compile_caller(tmp_path)
    res = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ok host" in res.stdout


@pytest.mark.skipif(not os.path.isdir(REF_INC), reason="reference headers not present")
def test_struct_layout_vs_reference_headers(tmp_path):
    ours = _layout_program(INC, tmp_path, "ours")
    ref = _layout_program(REF_INC, tmp_path, "ref")
    assert len(ours) == sum(1 + len(v) for v in LAYOUT.values()) + len(ENUMS)
    assert ours == ref


@pytest.mark.gpu
def test_c_caller_full(gpu, tmp_path):
    print("case: C caller, full", flush=True)
    sys.stdout.flush()
    exe = _compile_caller(tmp_path)
    res = subprocess.run([exe, "full"], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ok full" in res.stdout
