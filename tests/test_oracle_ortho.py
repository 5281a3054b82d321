"""Pins of the oracle's orthoQS restatement (SVD truncation, CPU only).

  * the singular values of every group are those of the dense matrix of the
    full tensor with legs 0..K-1 fused against the rest (union over groups,
    the remaining dense values zero);
  * A1 . A2 over the new leg gives back PSI when everything is kept, and the
    dense Eckart-Young best rank-Nk approximation times sfac when Nk states
    are kept across a gap;
  * A1 is an isometry: A1^dagger A1 = identity on the new leg;
  * the markSet rules (Nkeep, stol, Nkmin) and the itags of the new leg by
    cases whose expected result is written out literally.
"""
import numpy as np
import pytest

import oracle as O
from fullspace import _layout, _sectors, expand
from qspace_amd import QSpace, gen
from test_oracle_blockops import u1_rank4


def _dense(T):
    lay = [_layout(_sectors(T, k, None)) for k in range(T.rank)]
    return expand(T, lay), lay


def _psi(seed=3):
    return gen.u1_two_site_psi(seed=seed, Lq=(-4, -2, 0, 2, 4), Rq=(-4, -2, 0, 2, 4), dmin=2, dmax=7)


CASES = [("psi", 1), ("psi", 2), ("rank4", 2), ("rank4", 1), ("mps", 2), ("mps", 1)]


def _tensor(kind):
    if kind == "psi":
        return _psi()
    if kind == "rank4":
        return u1_rank4(5, nq=2, dmin=1, dmax=5)
    return gen.u1_mps(seed=9, Lq=(-4, -2, 0, 2, 4), Rq=(-3, -1, 1, 3), dmin=2, dmax=6)


@pytest.mark.parametrize("kind,K", CASES)
def test_singular_values_are_the_dense_ones(kind, K):
    P = _tensor(kind)
    A1, A2, info = O.ortho(P, K)
    F, _ = _dense(P)
    d1 = int(np.prod(F.shape[:K]))
    s = np.linalg.svd(F.reshape(d1, -1, order="F"), compute_uv=False)
    n = info["Ntot"]
    assert np.allclose(np.sort(info["svd"])[::-1], s[:n], rtol=1e-12, atol=1e-12 * s[0])
    assert np.all(s[n:] < 1e-12 * s[0])
    assert info["sv2tot"] == pytest.approx(float(np.sum(F * F)), rel=1e-13)


@pytest.mark.parametrize("kind,K", CASES)
def test_keep_all_reconstructs(kind, K):
    P = _tensor(kind)
    A1, A2, info = O.ortho(P, K)
    assert info["Nkeep"] == info["Ntot"] and info["sfac"] == 1.0 and info["svd2tr"] == 0.0
    assert A1.rank == K + 1 and A2.rank == P.rank - K + 1
    X = O.contract(A1, [K], A2, [P.rank - K])
    FP, lay = _dense(P)
    FX = expand(X, lay)
    assert np.abs(FX - FP).max() < 1e-12 * np.abs(FP).max()


@pytest.mark.parametrize("kind,K", CASES)
def test_u_is_an_isometry(kind, K):
    P = _tensor(kind)
    A1, _, _ = O.ortho(P, K, Nkeep=7)
    G = O.contract(A1, list(range(K)), A1, list(range(K)), conjA=1)
    for i, d in enumerate(G.data):
        assert G.Q[0][i, 0] == G.Q[1][i, 0]
        assert np.allclose(d, np.eye(d.shape[0]), atol=1e-13)


def test_truncation_is_the_best_low_rank_approximation():
    P = _psi(11)
    F, lay = _dense(P)
    M = F.reshape(F.shape[0] * F.shape[1], -1, order="F")
    U, s, Vt = np.linalg.svd(M, full_matrices=False)
    nk = next(k for k in range(6, len(s)) if s[k - 1] - s[k] > 1e-6 * s[0])
    A1, A2, info = O.ortho(P, 2, Nkeep=nk)
    assert info["Nkeep"] == nk
    sfac = np.sqrt(np.sum(s * s)) / np.sqrt(np.sum(s[:nk] ** 2))
    assert info["sfac"] == pytest.approx(sfac, rel=1e-13)
    assert info["svd2tr"] == pytest.approx(np.sum(s[nk:] ** 2) / np.sum(s * s), rel=1e-12)
    best = (U[:, :nk] * s[:nk]) @ Vt[:nk] * sfac
    X = O.contract(A1, [2], A2, [2])
    FX = expand(X, lay).reshape(M.shape, order="F")
    assert np.abs(FX - best).max() < 1e-12 * np.abs(best).max()


def test_mark_set_rules():
    E = np.array([0.1, 0.2, 0.3, 0.5, 0.5, 0.9, 1.5, 2.0])  # ascending
    f = O._mark_set_desc
    # Nkeep only: the top Nkeep
    m, nk = f(E, 0, 3, 0.0)
    assert nk == 3 and list(m) == [0, 0, 0, 0, 0, 1, 1, 1]
    # stol: everything above the highest value <= stol
    m, nk = f(E, 0, 100, 0.5)
    assert nk == 3 and list(m) == [0, 0, 0, 0, 0, 1, 1, 1]
    m, nk = f(E, 0, 100, 0.45)
    assert nk == 5 and list(m) == [0, 0, 0, 1, 1, 1, 1, 1]
    # stol capped by Nkeep
    m, nk = f(E, 0, 2, 0.25)
    assert nk == 2
    # Nkmin: the top Nkmin are kept whatever stol says
    m, nk = f(E, 6, 100, 0.95)
    assert nk == 6 and list(m) == [0, 0, 1, 1, 1, 1, 1, 1]
    # nothing below stol within the allowed range: all kept
    m, nk = f(E, 0, 100, 0.05)
    assert nk == 8 and all(m)
    m, nk = f(E, 8, 100, 3.0)
    assert nk == 8 and all(m)
    # Nkeep = 0 keeps nothing; Nkeep >= N without stol keeps all
    assert f(E, 0, 0, 0.0)[1] == 0
    assert f(E, 0, 8, 0.0)[1] == 8
    # every value <= stol and no Nkmin: nothing kept
    m, nk = f(E, 0, 100, 5.0)
    assert nk == 0 and not any(m)
    # ties keep the stable order: of two equal values the later one is above
    marks, nk = O._mark_set([np.array([0.5, 0.2]), np.array([0.5, 0.1])], 0, 1, 0.0)
    assert nk == 1 and list(marks[0]) == [0, 0] and list(marks[1]) == [1, 0]


def test_new_leg_itags():
    P = _psi()
    # K = r-1: the name (or the given one) and direction of the last leg
    A1, A2, _ = O.ortho(gen.u1_mps(seed=2, Lq=(-2, 0, 2), Rq=(-1, 1), dmin=2, dmax=3), 2)
    assert A1.itags == ["L", "s", "R*"] and A2.itags == ["R*", "R"]
    A1, A2, _ = O.ortho(gen.u1_mps(seed=2, Lq=(-2, 0, 2), Rq=(-1, 1), dmin=2, dmax=3), 2, itag="X")
    assert A1.itags == ["L", "s", "X*"] and A2.itags == ["R*", "X"]
    # K = 1: the first leg's name, opposite direction
    A1, A2, _ = O.ortho(P, 1)
    assert A1.itags == ["L", "L*"] and A2.itags == ["s1", "s2*", "R*", "L"]
    A1, A2, _ = O.ortho(P, 1, itag="Y*")
    assert A1.itags == ["L", "Y*"] and A2.itags == ["s1", "s2*", "R*", "Y"]
    # otherwise the given name, outgoing when the first leg is incoming
    A1, A2, _ = O.ortho(P, 2, itag="Z")
    assert A1.itags == ["L", "s1", "Z*"] and A2.itags == ["s2*", "R*", "Z"]
    # new-leg labels: the group's label sum
    for i in range(A1.nblk):
        assert A1.Q[2][i, 0] == A1.Q[0][i, 0] + A1.Q[1][i, 0]


def test_errors():
    P = _psi()
    with pytest.raises(O.OracleError) as e:
        O.ortho(P, 3)  # s2* among the fused (ingoing) legs
    assert e.value.code == 2
    with pytest.raises(O.OracleError):
        O.ortho(P, 0)
    with pytest.raises(O.OracleError):
        O.ortho(P, 4)
    Q2 = QSpace(P.Q, P.data, "A", ["L", "s1", "s2*", "R"])
    with pytest.raises(O.OracleError) as e:
        O.ortho(Q2, 2)  # mixed directions in the right group
    assert e.value.code == 2
    # K = r-1 with an incoming last leg: the identity-CG flip is not restated
    M = gen.u1_mps(seed=2, Lq=(-2, 0, 2), Rq=(-1, 1), dmin=2, dmax=3)
    M2 = QSpace(M.Q, M.data, "A", ["L", "s", "R"])
    with pytest.raises(O.OracleError) as e:
        O.ortho(M2, 2)
    assert e.value.code == 5


def u1_rank2(seed=4, qs=(-2, 0, 2), dmin=2, dmax=6):
    """block-diagonal U(1) matrix psi(L, R*), one block per charge"""
    rng = np.random.RandomState(seed)
    Q, data = [[], []], []
    for q in qs:
        d1, d2 = int(rng.randint(dmin, dmax + 1)), int(rng.randint(dmin, dmax + 1))
        Q[0].append([q])
        Q[1].append([q])
        data.append(rng.standard_normal((d1, d2)))
    return QSpace([np.array(x) for x in Q], data, "A", ["L", "R*"])


def test_nothing_kept():
    P = _psi()
    for kw in ({"Nkeep": 0}, {"Nkeep": 50, "stol": 1e3}):
        A1, A2, info = O.ortho(P, 2, **kw)
        assert A1.nblk == 0 and A2.nblk == 0 and info["Nkeep"] == 0
        assert info["svd2tr"] == 1.0 and info["sfac"] == 1.0
        assert A1.itags == ["L", "s1", "*"] and A2.itags == ["s2*", "R*", ""]


def test_rank_deficient_group():
    """a zero block makes its group rank deficient: the zero singular values
    are reported and never kept ahead of non-zero ones"""
    P = _psi(5)
    data = [np.zeros_like(d) if i == 3 else d for i, d in enumerate(P.data)]
    Z = QSpace(P.Q, data, P.qtype, P.itags)
    A1, A2, info = O.ortho(Z, 2, Nkeep=20)
    F, lay = _dense(Z)
    s = np.linalg.svd(F.reshape(F.shape[0] * F.shape[1], -1, order="F"), compute_uv=False)
    assert np.allclose(info["svd"][:20], s[:20], atol=1e-12 * s[0])
    assert info["Nkeep"] == 20


def test_rank2_block_diagonal():
    """rank-2 psi (K = 1 = r-1): the new leg takes the last leg's itag, A2 is
    a rank-2 block-diagonal matrix (its zero blocks are kept)"""
    P = u1_rank2()
    A1, A2, info = O.ortho(P, 1, Nkeep=5)
    assert A1.itags == ["L", "R*"] and A2.itags == ["R*", "R"]
    for i in range(A2.nblk):
        assert A2.Q[0][i, 0] == A2.Q[1][i, 0]
    X = O.contract(A1, [1], A2, [1])
    F, lay = _dense(P)
    U, s, Vt = np.linalg.svd(F)
    best = (U[:, :5] * s[:5]) @ Vt[:5] * (np.sqrt(np.sum(s * s)) / np.sqrt(np.sum(s[:5] ** 2)))
    assert np.abs(expand(X, lay) - best).max() < 1e-12 * np.abs(best).max()


def test_front_end_defaults():
    """orthoQS front-end defaults (MPS_ORTHO_1, Source/orthoQS.cc:137-170,
    get_Nkeep_estimate :39-56) on a hand-built rank-3 tensor whose leg dims
    (sums of block dims over distinct labels) are 5, 2 and 9"""
    import qspace_amd as Q
    Qs = [np.array([[0], [2], [0]]), np.array([[1], [1], [-1]]), np.array([[1], [3], [-1]])]
    data = [np.ones((2, 1, 4)), np.ones((3, 1, 2)), np.ones((2, 1, 3))]
    P = QSpace(Qs, data, "A", ["a", "b", "c*"])
    # leg dims: leg 0 {0: 2, 2: 3} = 5, leg 1 {1: 1, -1: 1} = 2, leg 2 {1: 4, 3: 2, -1: 3} = 9
    assert Q.ortho_args(P, 2) == (min(10, 9), 1e-8)
    assert Q.ortho_args(P, 1) == (min(5, 18), 1e-8)
    assert Q.ortho_args(P, 2, Nkeep=-3) == (9, 1e-8)
    assert Q.ortho_args(P, 2, Nkeep=3) == (3, 0.0)          # Nkeep > 0 without stol: stol = 0
    assert Q.ortho_args(P, 2, Nkeep=3, stol=0.25) == (3, 0.25)
    assert Q.ortho_args(P, 2, Nkeep=0) == (0, 1e-8)         # Nkeep = 0 is not > 0
    S = QSpace(Qs, data, "SU2", ["a", "b", "c*"])
    assert Q.ortho_args(S, 2) == (max(10, 9), 1e-8)          # non-abelian: the larger side
    assert Q.ortho_args(S, 1) == (max(5, 18), 1e-8)
