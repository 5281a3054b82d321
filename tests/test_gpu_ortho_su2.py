"""SU(2) orthoQS (K = r-1) on the device vs the CPU oracle (GPU tests), to
the bar of the abelian truncation tests (structure, weights, kept count and
flag bit-exact; singular values within 1e-12 of the largest; vectors up to
a sign per column; the product A1 . A2 and the isometry gauge-free), plus
the multiplet dimension of every singular value bit-exact."""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from test_gpu_ortho import _check
from test_oracle_ortho_su2 import _mixed_mps, _su2_matrix, _tensor, _weighted

pytestmark = pytest.mark.gpu


def _big(kind):
    if kind == "mps_full":  # group matrices up to 128 x 64: in LDS
        return gen.su2_mps(seed=1002)[0]
    if kind == "mps_wide":  # group matrices beyond the LDS: block Jacobi steps
        return _weighted(gen.su2_mps(seed=1003, tS=tuple(range(0, 9)), dmin=80, dmax=150)[0])
    raise KeyError(kind)


CASES = [("mps", {}), ("mps", {"Nkeep": 4}), ("mixed", {}), ("mixed", {"Nkeep": 6, "stol": 0.5}),
         ("matrix", {}), ("matrix", {"Nkeep": 2}), ("matrix_op", {"Nkeep": 3}), ("weighted", {"Nkeep": 3}),
         ("weighted", {"stol": 1.0, "Nkmin": 2})]


@pytest.mark.parametrize("kind,kw", CASES)
def test_su2_ortho_matches_oracle(gpu, kind, kw):
    P = _tensor(kind)
    _check(P, P.rank - 1, **kw)
    _, _, info = Q.orthoQS(P, P.rank - 1, **kw)
    _, _, rinfo = O.ortho(P, P.rank - 1, **kw)
    assert np.array_equal(info["svd_qdim"], rinfo["svd_qdim"])


@pytest.mark.parametrize("kind,kw", [("mps_full", {"Nkeep": 40}), ("mps_wide", {"Nkeep": 60, "stol": 1e-3})])
def test_su2_ortho_full_size(gpu, kind, kw):
    P = _big(kind)
    _check(P, P.rank - 1, **kw)


def test_su2_ortho_unsupported(gpu):
    P = _tensor("mps")
    with pytest.raises(Q.QsxError) as e:
        Q.orthoQS(P, 1)
    assert e.value.code == 5
