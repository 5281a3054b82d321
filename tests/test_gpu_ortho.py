"""orthoQS (block SVD truncation) on the device vs the CPU oracle (GPU tests).

Bar: the structure of both results (label records, block order, dims,
itags), the kept count, Ntot and flag bit-exact; singular values, svd2tr,
sv2tot and sfac within 1e-12 (relative to the largest singular value /
themselves); singular vectors up to the sign of each column (the SVD's own
gauge), within 1e-9, and gauge-free: A1 . A2 equals the oracle's product
within 1e-12, A1 is an isometry within 1e-12.
"""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from helpers import assert_same_structure, oracle_ortho
from test_oracle_blockops import u1_rank4

pytestmark = pytest.mark.gpu


def _tensor(kind):
    if kind == "psi":
        return gen.u1_two_site_psi(seed=3, Lq=(-4, -2, 0, 2, 4), Rq=(-4, -2, 0, 2, 4), dmin=2, dmax=7)
    if kind == "rank4":
        return u1_rank4(5, nq=2, dmin=1, dmax=5)
    if kind == "mps":
        return gen.u1_mps(seed=9, Lq=(-4, -2, 0, 2, 4), Rq=(-3, -1, 1, 3), dmin=2, dmax=6)
    if kind == "psi_full":  # every group matrix in LDS (dims up to 2 x 32)
        return gen.u1_two_site_psi(seed=1007)
    if kind == "psi_wide":  # group matrices beyond the LDS capacity
        return gen.u1_two_site_psi(seed=1008, Lq=(-4, -2, 0, 2, 4), Rq=(-4, -2, 0, 2, 4), dmin=60, dmax=110)
    raise KeyError(kind)


def _group_mats(T):
    """per new-leg label: the blocks of T stacked along their fused leading
    legs (T's last leg is the new one) -> {label: matrix}"""
    out = {}
    dims = T.dims()
    for i, d in enumerate(T.data):
        key = tuple(T.Q[T.rank - 1][i])
        out.setdefault(key, []).append(np.asarray(d).reshape(-1, int(dims[i, -1]), order="F"))
    return {k: np.vstack(v) for k, v in out.items()}


def _check(P, K, **kw):
    A1, A2, info = Q.orthoQS(P, K, **kw)
    R1, R2, rinfo = oracle_ortho(P, K, **kw)
    print(f"orthoQS K={K} {kw} blocks={P.nblk} groups={info['ngroups']} sweeps={info['sweeps']} "
          f"keep={info['Nkeep']}/{info['Ntot']}", flush=True)
    assert_same_structure(A1, R1)
    assert_same_structure(A2, R2)
    for k in ("Nkeep", "Ntot", "flag"):
        assert info[k] == rinfo[k], k
    smax = float(rinfo["svd"][0])
    assert np.abs(info["svd"] - rinfo["svd"]).max() <= 1e-12 * smax
    for k in ("sfac", "sv2tot"):
        assert info[k] == pytest.approx(rinfo[k], rel=1e-12), k
    assert info["svd2tr"] == pytest.approx(rinfo["svd2tr"], rel=1e-9, abs=1e-15)
    # singular vectors up to a sign per column
    U, UR = _group_mats(A1), _group_mats(R1)
    V, VR = _group_mats(A2), _group_mats(R2)
    for key in UR:
        sgn = np.sign(np.sum(U[key] * UR[key], axis=0))
        assert np.all(sgn != 0)
        assert np.abs(U[key] * sgn - UR[key]).max() < 1e-9
        if key in VR:
            assert np.abs(V[key] * sgn - VR[key]).max() < 1e-9 * smax
    # gauge-free: the product and the isometry
    if A1.nblk and A2.nblk:
        X = O.contract(A1, [K], A2, [P.rank - K])
        XR = O.contract(R1, [K], R2, [P.rank - K])
        assert_same_structure(X, XR)
        num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(X.data, XR.data))
        den = sum(float(np.sum(b * b)) for b in XR.data)
        assert np.sqrt(num / den) < 1e-12
        G = O.contract(A1, list(range(K)), A1, list(range(K)), conjA=1)
        for d in G.data:
            assert np.abs(d - np.eye(d.shape[0])).max() < 1e-12
    return A1, A2, info


CASES = [("psi", 2, {}), ("psi", 1, {}), ("psi", 2, {"Nkeep": 9}), ("psi", 2, {"Nkeep": 100, "stol": 0.5}),
         ("psi", 2, {"Nkeep": 100, "stol": 0.5, "Nkmin": 14, "itag": "X"}), ("rank4", 2, {"Nkeep": 12}),
         ("rank4", 1, {}), ("mps", 2, {"Nkeep": 10}), ("mps", 1, {"stol": 0.3})]


@pytest.mark.parametrize("kind,K,kw", CASES)
def test_ortho_matches_oracle(gpu, kind, K, kw):
    _check(_tensor(kind), K, **kw)


@pytest.mark.parametrize("kind", ["psi_full", "psi_wide"])
def test_ortho_full_size(gpu, kind):
    """the two-site psi at full size (LDS path) and with group matrices
    beyond the LDS (the in-place path): parity, and A1 . A2 = PSI when all
    states are kept"""
    P = _tensor(kind)
    _check(P, 2, Nkeep=400)
    A1, A2, info = Q.orthoQS(P, 2)
    assert info["Nkeep"] == info["Ntot"] and info["sfac"] == 1.0
    X = O.contract(A1, [2], A2, [2])
    got = {tuple(X.QIDX[i]): d for i, d in enumerate(X.data)}
    assert sorted(got) == sorted(tuple(r) for r in P.QIDX)
    num = sum(float(np.sum((got[tuple(P.QIDX[i])] - d) ** 2)) for i, d in enumerate(P.data))
    den = sum(float(np.sum(d * d)) for d in P.data)
    assert np.sqrt(num / den) < 1e-12


def test_ortho_unsupported_flip(gpu):
    M = gen.u1_mps(seed=2, Lq=(-2, 0, 2), Rq=(-1, 1), dmin=2, dmax=3)
    M2 = Q.QSpace(M.Q, M.data, "A", ["L", "s", "R"])
    with pytest.raises(Q.QsxError) as e:
        Q.orthoQS(M2, 2)
    assert e.value.code == 5


def test_ortho_edge_cases(gpu):
    """nothing kept (Nkeep = 0; stol above every value), a group with a zero
    block, a rank-2 block-diagonal psi (K = 1 = r-1)"""
    from test_oracle_ortho import u1_rank2
    P = _tensor("psi")
    for kw in ({"Nkeep": 0}, {"Nkeep": 50, "stol": 1e3}):
        A1, A2, info = _check(P, 2, **kw)
        assert A1.nblk == 0 and A2.nblk == 0 and info["Nkeep"] == 0
    Z = gen.u1_two_site_psi(seed=5, Lq=(-4, -2, 0, 2, 4), Rq=(-4, -2, 0, 2, 4), dmin=2, dmax=7)
    Z = Q.QSpace(Z.Q, [np.zeros_like(d) if i == 3 else d for i, d in enumerate(Z.data)], Z.qtype, Z.itags)
    _check(Z, 2, Nkeep=20)
    _check(u1_rank2(), 1, Nkeep=5)
    _check(u1_rank2(seed=6, qs=(-4, -2, 0, 2, 4), dmin=20, dmax=150), 1, Nkeep=200, stol=1e-3)


def test_ortho_rank_deficient_groups(gpu):
    """group matrices of exactly deficient rank (every block of a two-site
    psi a product of thin factors, several blocks per group): the zero
    singular values come out at rounding level, the sweeps end (a column of
    rounding noise is not rotated against a larger one), and the truncated
    and the full products agree with the oracle"""
    rng = np.random.RandomState(17)
    P = gen.u1_two_site_psi(seed=6, Lq=(-4, -2, 0, 2, 4), Rq=(-4, -2, 0, 2, 4), dmin=3, dmax=7)
    data = []
    for d in P.data:
        s = d.shape
        L = rng.standard_normal((s[0] * s[1], 2))
        R = rng.standard_normal((2, s[2] * s[3]))
        data.append(np.asfortranarray((L @ R).reshape(s, order="F")))
    D = Q.QSpace(P.Q, data, P.qtype, P.itags)
    for kw in ({}, {"Nkeep": 6}):
        A1, A2, info = Q.orthoQS(D, 2, **kw)
        R1, R2, rinfo = oracle_ortho(D, 2, **kw)
        assert (info["Nkeep"], info["Ntot"]) == (rinfo["Nkeep"], rinfo["Ntot"])
        sv, rsv = np.asarray(info["svd"]), np.asarray(rinfo["svd"])
        assert np.abs(sv - rsv).max() <= 1e-12 * rsv.max()
        assert (rsv < 1e-13 * rsv.max()).sum() > 0  # the draw is rank-deficient
        X = O.contract(A1, [2], A2, [2], skip_zero=0)
        XR = O.contract(R1, [2], R2, [2], skip_zero=0)
        assert_same_structure(X, XR)
        num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(X.data, XR.data))
        den = sum(float(np.sum(b * b)) for b in XR.data)
        assert np.sqrt(num / den) < 1e-12
