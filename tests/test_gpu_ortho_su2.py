"""SU(2) orthoQS (K = r-1, and K = 1 of a rank-3 tensor) on the device vs
the CPU oracle (GPU tests), to the bar of the abelian truncation tests (structure, weights, kept count and
flag bit-exact; singular values within 1e-12 of the largest; vectors up to
a sign per column; the product A1 . A2 and the isometry gauge-free), plus
the multiplet dimension of every singular value bit-exact."""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from helpers import oracle_ortho
from test_gpu_ortho import _check
from test_oracle_ortho_su2 import _mixed_mps, _su2_matrix, _tensor, _weighted

pytestmark = pytest.mark.gpu


def _big(kind):
    if kind == "mps_full":  # group matrices up to 128 x 64: in LDS
        return gen.su2_mps(seed=1002)[0]
    if kind == "mps_wide":  # group matrices beyond the LDS: block Jacobi steps
        return _weighted(gen.su2_mps(seed=1003, tS=tuple(range(0, 9)), dmin=80, dmax=150)[0])
    raise KeyError(kind)


CASES = [("mps", 2, {}), ("mps", 2, {"Nkeep": 4}), ("mixed", 2, {}), ("mixed", 2, {"Nkeep": 6, "stol": 0.5}),
         ("matrix", 1, {}), ("matrix", 1, {"Nkeep": 2}), ("matrix_op", 1, {"Nkeep": 3}),
         ("weighted", 2, {"Nkeep": 3}), ("weighted", 2, {"stol": 1.0, "Nkmin": 2}),
         ("mps", 1, {}), ("mps", 1, {"Nkeep": 5}), ("mixed", 1, {"Nkeep": 5, "stol": 0.3}), ("weighted", 1, {})]


@pytest.mark.parametrize("kind,K,kw", CASES)
def test_su2_ortho_matches_oracle(gpu, kind, K, kw):
    P = _tensor(kind)
    _check(P, K, **kw)
    _, _, info = Q.orthoQS(P, K, **kw)
    _, _, rinfo = oracle_ortho(P, K, **kw)
    assert np.array_equal(info["svd_qdim"], rinfo["svd_qdim"])


@pytest.mark.parametrize("kind,K,kw", [("mps_full", 2, {"Nkeep": 40}), ("mps_wide", 2, {"Nkeep": 60, "stol": 1e-3}),
                                       ("mps_full", 1, {"Nkeep": 40}), ("mps_wide", 1, {"Nkeep": 60})])
def test_su2_ortho_full_size(gpu, kind, K, kw):
    P = _big(kind)
    _check(P, K, **kw)


def test_su2_ortho_unsupported(gpu):
    A = Q.QSpace([np.array([[0], [1]], dtype=np.int32) for _ in range(4)], [np.ones((2, 1, 1, 2))] * 2, "SU2",
                 ["a", "b", "c*", "d*"])
    with pytest.raises(Q.QsxError) as e:
        Q.orthoQS(A, 2)
    assert e.value.code == 5
