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
    """rank 4 splits other than K = 2 are refused"""
    A = gen.su2_om_tensor(5, [[(0, 2), (1, 1)]] * 4, ["a", "b", "c*", "d*"])
    B = gen.su2_om_tensor(6, [[(0, 2), (1, 1)]] * 4, ["a", "b", "c", "d*"])
    for T, K in ((A, 1), (B, 3)):
        with pytest.raises(Q.QsxError) as e:
            Q.orthoQS(T, K)
        assert e.value.code == 5


def _two_site(seed, labels, dmin, dmax, tags=("L", "s1", "s2*", "R*")):
    rng = np.random.RandomState(seed)
    L = [(t, int(rng.randint(dmin, dmax + 1))) for t in labels]
    R = [(t, int(rng.randint(dmin, dmax + 1))) for t in labels]
    return gen.su2_om_tensor(seed, [L, [(1, 1)], [(1, 1)], R], list(tags))


@pytest.mark.parametrize("kw", [{"Nkeep": 1000}, {"Nkeep": 12}, {"Nkeep": 30, "stol": 1e-2}])
def test_su2_two_site_rank4_truncation(gpu, kw):
    """K = 2 of the rank-4 SU(2) two-site wave function (legs with outer
    multiplicity; mpsortho.cc:1031-1078 after eigs2SiteQS.cc:481-484), in the
    full M-resolved space: the singular values, each repeated by its
    multiplet dimension, are the dense singular values of Full(PSI) split
    (L, s1 | s2, R); A1 is an isometry; A1.A2 = PSI when all are kept and
    = sfac times the projection onto A1's columns after a truncation"""
    from test_oracle_ortho_su2 import _full_parts
    P = _two_site(61, range(0, 7), 2, 6)
    assert max(P.om()) > 1
    A1, A2, info = Q.orthoQS(P, 2, itag="X", **kw)
    print(f"rank-4 K=2 {kw}: Ntot={info['Ntot']} Nkeep={info['Nkeep']}", flush=True)
    assert A1.itags == ["L", "s1", "X*"] and A2.itags == ["s2*", "R*", "X"]
    FP, F1, F2 = _full_parts(P, A1, A2, 2)
    d1 = int(np.prod(FP.shape[:2]))
    s = np.linalg.svd(FP.reshape(d1, -1), compute_uv=False)
    full = np.sort(np.repeat(info["svd"], info["svd_qdim"].astype(int)))[::-1]
    assert np.allclose(full, s[:len(full)], rtol=1e-12, atol=1e-12 * s[0])
    assert np.all(s[len(full):] < 1e-12 * s[0])
    assert info["sv2tot"] == pytest.approx(float(np.sum(FP * FP)), rel=1e-12)
    M1 = F1.reshape(-1, F1.shape[-1])
    assert np.allclose(M1.T @ M1, np.eye(M1.shape[1]), atol=1e-12)
    X = np.tensordot(F1, F2, axes=([2], [2]))
    proj = (M1 @ (M1.T @ FP.reshape(d1, -1))).reshape(FP.shape)
    assert np.allclose(X, info["sfac"] * proj, atol=1e-12 * np.abs(FP).max())
    if kw["Nkeep"] >= info["Ntot"]:
        assert info["Nkeep"] == info["Ntot"] and info["sfac"] == 1.0
        assert np.allclose(X, FP, atol=1e-12 * np.abs(FP).max())
    else:
        assert info["Nkeep"] <= kw["Nkeep"]
        assert M1.shape[1] == int(np.sum(info["svd_qdim"][:info["Nkeep"]]))


def test_su2_two_site_rank4_truncation_large(gpu):
    """the same at 2S <= 12 with dims up to 24 per multiplet: A1.A2 = PSI
    block by block (contracting the results on the device, compared with
    PSI to 1e-12) and the kept weight"""
    P = _two_site(62, range(0, 13), 8, 24)
    A1, A2, info = Q.orthoQS(P, 2, Nkeep=100000, itag="X")
    assert info["Nkeep"] == info["Ntot"]
    B = Q.contractQS(A1, "3", A2, "3")
    from test_gpu_om import assert_same_by_record
    assert_same_by_record(B, P)
