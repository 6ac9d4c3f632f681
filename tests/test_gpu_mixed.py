"""Mixed-precision products: off-diagonal factors split into 1..4 BF16
slices on MFMA with exact FP32 accumulation, diagonal leaves FP64, vs the
FP64 dense product of the same tree.

Tolerance per slice count (relative, Frobenius): each slice carries 8 bits
of the chunk maximum, so 1 slice 3e-2, 2 slices 5e-4, 3 slices 2e-6, 4
slices 1e-8 (the workload target).  Trees cover
ragged and empty leaves, leaves longer than one 128-row chunk and ranks
spanning two 128-rank chunks."""
import sys

import numpy as np
import pytest

from hmat_lib_amd import capi
from hmat_lib_amd import tree as T
from tests import helpers as H

pytestmark = pytest.mark.gpu
TOLS = {1: 3e-2, 2: 5e-4, 3: 2e-6, 4: 1e-8}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


CASES = [
    ("uniform h3", [128] * 8, lambda l, j: 5 + 3 * l),
    ("ragged zeros h3", [200, 0, 37, 129, 64, 300, 1, 90], lambda l, j: 1 + (5 * l + j) % 23),
    ("rank 150 h2", [260, 250, 240, 270], lambda l, j: 150 if l == 1 else 70),
]


@pytest.mark.parametrize("name,leaves,ranks", CASES, ids=[c[0] for c in CASES])
def test_mixed_products_vs_dense(gpu, name, leaves, ranks):
    from hmat_lib_amd import device as D
    rng = np.random.default_rng(len(leaves) + sum(leaves))
    ft = H.random_flat_tree(rng, leaves, ranks)
    lib = capi.hmat()
    keep = T.Keep()
    t = T.fill_tree(lib, ft, keep)
    try:
        dense = ft.to_dense()
        n = ft.n
        dt = D.DeviceTree.upload(t.contents.root, ft.height)
        for R in (3, 64, 100):
            x = np.asfortranarray(rng.standard_normal((n, R)))
            ref = dense @ x
            dx = D.DeviceArray.from_numpy(x)
            dy = D.DeviceArray((n, R))
            for S in (1, 2, 3, 4):
                _say(f"case: mixed {name} R={R} slices={S}")
                dt.set_precision(S)
                dt.matmat(dx, dy)
                dt.matmat(dx, dy)  # cached program, same result
                err = H.rel_err(dy.numpy(), ref)
                assert err < TOLS[S], (S, err)
                # transposed products stay FP64
                dt.matmat(dx, dy, trans=1)
                assert H.rel_err(dy.numpy(), dense.T @ x) < 1e-12
            dt.set_precision(0)
            dt.matmat(dx, dy)
            assert H.rel_err(dy.numpy(), ref) < 1e-12
        dt.free()
    finally:
        T.detach_and_free(lib, t)
