"""Pins of the oracle's SU(2) orthoQS (K = r-1, and K = 1 of a rank-3
tensor; CPU only), all in the full M-resolved space, where they hold
whatever the CGT conventions are:

  * the singular values, each repeated by its multiplet dimension, are the
    dense singular values of Full(PSI) split (legs 0..K-1 | legs K..r-1), and
    sv2tot = |Full(PSI)|^2 (drops the sqrt(qdim) of the block scale or the
    qdim weight of the norms if either is wrong);
  * Full(A1) is an isometry over the new leg;
  * Full(A1) . Full(A2) = Full(PSI) when everything is kept, and = sfac
    times the projection of Full(PSI) onto the kept columns of Full(A1)
    after a truncation, with |Full(A1) . Full(A2)| = |Full(PSI)|;
  * blocks carrying weights other than normExt (the block scale uses cgw).
"""
import itertools

import numpy as np
import pytest

import oracle as O
from fullspace import _layout, _sectors, expand
from qspace_amd import QSpace, gen


def _su2_j(T):
    return [j for j, s in enumerate(O._syms(T)) if s == "SU2"][0]


def _mixed_mps(seed=5):
    """A,SU2 rank-3 A(L, s, R*): U(1) charge q_L + q_s = q_R, SU(2) triangle"""
    rng = np.random.RandomState(seed)
    legL = {(q, t): int(rng.randint(1, 5)) for q in (-1, 0, 1) for t in (0, 1, 2, 3)}
    legs = {(1, 1): 1, (-1, 1): 1, (0, 0): 1}
    legR = {(q, t): int(rng.randint(1, 5)) for q in (-2, -1, 0, 1, 2) for t in (0, 1, 2, 3, 4)}
    Q, data = [[], [], []], []
    for a, b, c in itertools.product(sorted(legL), sorted(legs), sorted(legR)):
        if a[0] + b[0] != c[0] or (a[1] + b[1] + c[1]) % 2 or not abs(a[1] - b[1]) <= c[1] <= a[1] + b[1]:
            continue
        if rng.rand() < 0.2:
            continue
        for k, x in enumerate((a, b, c)):
            Q[k].append(list(x))
        data.append(np.asfortranarray(rng.standard_normal((legL[a], legs[b], legR[c]))))
    return QSpace([np.array(q, dtype=np.int32) for q in Q], data, "A,SU2", ["L", "s", "R*"])


def _su2_matrix(seed=6, dirs=(1, -1)):
    rng = np.random.RandomState(seed)
    Q, data = [[], []], []
    for t in (0, 1, 2, 3, 4):
        d1, d2 = int(rng.randint(1, 6)), int(rng.randint(1, 6))
        Q[0].append([t]); Q[1].append([t])
        data.append(np.asfortranarray(rng.standard_normal((d1, d2))))
    tags = ["a" + ("*" if dirs[0] < 0 else ""), "b" + ("*" if dirs[1] < 0 else "")]
    return QSpace([np.array(q, dtype=np.int32) for q in Q], data, "SU2", tags)


def _weighted(T, seed=7):
    """the same tensor with cgw = f * normExt and data / f (Full unchanged)"""
    rng = np.random.RandomState(seed)
    j = _su2_j(T)
    dirs = O._tag_dirs(T.itags, T.rank)
    cg, data = [], []
    for i, d in enumerate(T.data):
        f = float(rng.choice([0.5, 2.0, -1.0]))
        w = [1.0] * T.qdim
        w[j] = f * O.norm_ext([int(T.Q[k][i, j]) for k in range(T.rank)], dirs)
        cg.append(w)
        data.append(np.asfortranarray(d / f))
    return QSpace([q.copy() for q in T.Q], data, T.qtype, list(T.itags), np.array(cg))


def _tensor(kind):
    if kind == "mps":
        return gen.su2_mps(seed=11, tS=tuple(range(0, 7)), dmin=2, dmax=6)[0]
    if kind == "mixed":
        return _mixed_mps()
    if kind == "matrix":
        return _su2_matrix()
    if kind == "matrix_op":
        return _su2_matrix(dirs=(-1, 1))
    return _weighted(gen.su2_mps(seed=12, tS=tuple(range(0, 7)), dmin=2, dmax=6)[0])


KINDS = [("mps", 2), ("mixed", 2), ("matrix", 1), ("matrix_op", 1), ("weighted", 2), ("mps", 1), ("mixed", 1),
         ("weighted", 1)]


def _full_parts(P, A1, A2, K):
    j = _su2_j(P)
    r = P.rank
    lp = [_layout(_sectors(P, k, j)) for k in range(r)]
    ln = _layout(_sectors(A1, K, j))
    FP = expand(P, lp)
    F1 = expand(A1, lp[:K] + [ln])
    F2 = expand(A2, lp[K:] + [ln])
    return FP, F1, F2


@pytest.mark.parametrize("kind,K", KINDS)
def test_singular_values_and_norm(kind, K):
    P = _tensor(kind)
    r = P.rank
    A1, A2, info = O.ortho(P, K)
    FP, _, _ = _full_parts(P, A1, A2, K)
    d1 = int(np.prod(FP.shape[:K]))
    s = np.linalg.svd(FP.reshape(d1, -1), compute_uv=False)
    rep = np.sort(np.repeat(info["svd"], info["svd_qdim"].astype(int)))[::-1]
    assert len(rep) <= len(s)
    assert np.allclose(rep, s[:len(rep)], rtol=1e-12, atol=1e-12 * s[0])
    assert np.all(s[len(rep):] < 1e-12 * s[0])
    assert info["sv2tot"] == pytest.approx(float(np.sum(FP * FP)), rel=1e-13)
    # records sorted descending by (S, qdim)
    rec = list(zip(info["svd"], info["svd_qdim"]))
    assert rec == sorted(rec, reverse=True)


@pytest.mark.parametrize("kind,K", KINDS)
def test_keep_all_reconstructs_with_an_isometry(kind, K):
    P = _tensor(kind)
    r = P.rank
    A1, A2, info = O.ortho(P, K)
    assert info["Nkeep"] == info["Ntot"] and info["sfac"] == 1.0
    FP, F1, F2 = _full_parts(P, A1, A2, K)
    M1 = F1.reshape(-1, F1.shape[-1])
    assert np.allclose(M1.T @ M1, np.eye(M1.shape[1]), atol=1e-12)
    X = np.tensordot(F1, F2, axes=([K], [r - K]))
    assert np.allclose(X, FP, atol=1e-12 * np.abs(FP).max())
    # output weights are normExt (the default) on every SU(2) block
    j = _su2_j(P)
    for T in (A1, A2):
        dd = O._tag_dirs(T.itags, T.rank)
        for i in range(T.nblk):
            ne = O.norm_ext([int(T.Q[k][i, j]) for k in range(T.rank)], dd)
            assert T.info["cgw"][i, j] == ne


@pytest.mark.parametrize("kind,K,nkeep", [("mps", 2, 4), ("mixed", 2, 6), ("matrix", 1, 2), ("weighted", 2, 3),
                                           ("mps", 1, 4), ("mixed", 1, 5), ("weighted", 1, 3)])
def test_truncation_is_the_scaled_projection(kind, K, nkeep):
    P = _tensor(kind)
    r = P.rank
    A1, A2, info = O.ortho(P, K, Nkeep=nkeep)
    assert info["Nkeep"] == nkeep < info["Ntot"]
    FP, F1, F2 = _full_parts(P, A1, A2, K)
    M1 = F1.reshape(-1, F1.shape[-1])
    assert np.allclose(M1.T @ M1, np.eye(M1.shape[1]), atol=1e-12)
    X = np.tensordot(F1, F2, axes=([K], [r - K]))
    d1 = int(np.prod(FP.shape[:K]))
    proj = (M1 @ (M1.T @ FP.reshape(d1, -1))).reshape(FP.shape)
    assert np.allclose(X, info["sfac"] * proj, atol=1e-12 * np.abs(FP).max())
    assert np.sum(X * X) == pytest.approx(np.sum(FP * FP), rel=1e-12)
    # the discarded weight counts each multiplet qdim times
    s, d = info["svd"], info["svd_qdim"]
    assert info["svd2tr"] == pytest.approx(float(np.sum(s[nkeep:] ** 2 * d[nkeep:])) / info["sv2tot"], rel=1e-12)
    assert M1.shape[1] == int(np.sum(d[:nkeep]))


def test_unsupported_su2_splits():
    # rank 4 (outer multiplicity of the fused legs): not restated
    T4 = QSpace([np.array([[0], [1]], dtype=np.int32) for _ in range(4)], [np.ones((2, 1, 1, 2))] * 2, "SU2",
                ["a", "b", "c*", "d*"])
    for K in (1, 2, 3):
        with pytest.raises(O.OracleError) as e:
            O.ortho(T4, K)
        assert e.value.code == 5


@pytest.mark.parametrize("seed", range(8))
def test_k1_split_any_directions(seed):
    """K = 1 of rank-3 SU(2) tensors with random leg directions: A1.A2 =
    PSI in the full space and A1 an isometry (the permutation sign of A2's
    CGT differs between direction patterns)"""
    from test_gpu_random import _su2_leg, _su2_tensor
    rng = np.random.RandomState(500 + seed)
    dirs = [1 if rng.rand() < 0.5 else -1 for _ in range(3)]
    P = _su2_tensor(rng, [_su2_leg(rng) for _ in range(3)], dirs,
                    [f"x{i}" + ("*" if d < 0 else "") for i, d in enumerate(dirs)])
    if not P.nblk:
        pytest.skip("empty draw")
    A1, A2, info = O.ortho(P, 1)
    FP, F1, F2 = _full_parts(P, A1, A2, 1)
    M1 = F1.reshape(-1, F1.shape[-1])
    assert np.allclose(M1.T @ M1, np.eye(M1.shape[1]), atol=1e-12)
    X = np.tensordot(F1, F2, axes=([1], [2]))
    assert np.allclose(X, FP, atol=1e-12 * np.abs(FP).max())
