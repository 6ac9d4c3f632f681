"""Build recipes: the product libraries (gfx950) and the test oracle."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _make(path, jobs=8, extra=()):
    cmd = ["make", "-C", path, f"-j{jobs}", *extra]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd)} failed:\n{res.stdout[-4000:]}\n{res.stderr[-4000:]}")
    return res.stdout


def build_product(jobs=8, force=True):
    """Compile libhmat_hip.so / libhmat_lib.so / libhmat_lib_test.so."""
    csrc = os.path.join(HERE, "csrc")
    if force:
        _make(csrc, jobs, ("clean",))
    return _make(csrc, jobs)


def build_oracle(jobs=4):
    """Compile the oracle (and the reference build when its sources exist)."""
    return _make(os.path.join(ROOT, "oracle"), jobs, ("all",))
