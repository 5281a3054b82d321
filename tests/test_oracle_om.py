"""Outer multiplicity (SU(2) CGTs of rank > 3) in the CPU oracle and the
library's host coefficient engine (CPU only).

Pins, independent of the basis formulas:
  * the oracle's basis tensors are orthonormal, invariant under the total
    SU(2) generators (the dual representation on outgoing legs) and complete:
    their number is the dimension of the invariant subspace of the
    brute-force generator matrices and the combinatorial count of spin-0
    couplings; for rank <= 3 the single basis tensor is the oracle's block
    CGT (pinned in test_oracle.py);
  * the oracle's block-wise contraction with outer multiplicity equals the
    full M-resolved space tensordot (convention-independent);
  * the library's basis (exact Racah coefficients, sparse recursion) and its
    matrix-valued X-symbols equal the oracle's (lowering-operator tables,
    dense contractions) -- the same stated convention, built independently.
"""
import numpy as np
import pytest

import om as OM
import oracle as O
import qspace_amd as Q
from fullspace import check_contraction
from qspace_amd import gen


def _gens(tj):
    d = tj + 1
    m = np.array([tj / 2 - i for i in range(d)])
    j = tj / 2
    Jp = np.zeros((d, d))
    for i in range(1, d):
        Jp[i - 1, i] = np.sqrt(j * (j + 1) - m[i] * (m[i] + 1))
    return np.diag(m), Jp, Jp.T


def _generator_matrices(tj, dirs):
    n = int(np.prod([t + 1 for t in tj]))
    mats = []
    for g in range(3):
        M = np.zeros((n, n))
        for k, (t, d) in enumerate(zip(tj, dirs)):
            G = _gens(t)[g]
            if d < 0:
                G = -G.T
            ops = [np.eye(x + 1) for x in tj]
            ops[k] = G
            K = np.ones((1, 1))
            for o in reversed(ops):  # column-major: leg 0 fastest
                K = np.kron(K, o)
            M += K
        mats.append(M)
    return mats


CASES = [([2, 2, 2, 2], [1, 1, -1, -1]), ([1, 1, 1, 1], [1, 1, 1, -1]), ([2, 1, 1, 2], [1, 1, 1, 1]),
         ([4, 2, 3, 1], [-1, 1, 1, -1]), ([2, 2, 2, 2, 2], [1, 1, 1, -1, -1]), ([3, 1, 2, 2], [1, -1, 1, -1]),
         ([1, 1, 2], [1, 1, -1]), ([3, 3], [1, 1]), ([2, 4, 2, 2], [-1, -1, -1, 1]), ([0, 2, 2, 0], [1, 1, -1, -1])]


@pytest.mark.parametrize("tj,dirs", CASES)
def test_oracle_basis_orthonormal_invariant_complete(tj, dirs):
    B = OM.om_basis(tj, dirs)
    F = B.reshape(len(B), -1, order="C")
    assert np.abs(F @ F.T - np.eye(len(B))).max() < 1e-13
    mats = _generator_matrices(tj, dirs)
    for b in B:
        v = b.ravel(order="F")
        for M in mats:
            assert np.abs(M @ v).max() < 1e-12
    sv = np.linalg.svd(np.vstack(mats), compute_uv=False)
    null = int(np.sum(sv < 1e-9)) + max(0, mats[0].shape[1] - len(sv))
    assert len(B) == null == OM.om_count(tj)


def test_oracle_basis_rank3_is_block_cgt():
    rng = np.random.RandomState(4)
    seen = 0
    for _ in range(400):
        r = rng.randint(2, 4)
        tj = list(rng.randint(0, 6, size=r))
        dirs = list(rng.choice([1, -1], size=r))
        if OM.om_count(tj) == 0:
            continue
        B = OM.om_basis(tj, dirs)
        assert len(B) == 1
        assert np.abs(B[0] - O.block_cgt(tj, dirs)).max() < 1e-13
        seen += 1
    assert seen > 40


def test_library_basis_and_x_equal_oracle():
    rng = np.random.RandomState(5)
    n = 0
    for _ in range(150):
        r = rng.randint(2, 6)
        tj = list(rng.randint(0, 5, size=r))
        dirs = list(rng.choice([1, -1], size=r))
        Bo, Bp = OM.om_basis(tj, dirs), Q.su2_om_basis(tj, dirs)
        assert Bo.shape == Bp.shape
        if len(Bo):
            assert np.abs(Bo - Bp).max() < 1e-13
            n += 1
    assert n > 40
    xcases = [([2, 2, 2], [1, 1, -1], [2, 2, 2, 2], [1, -1, 1, -1], [2], [0], None),
              ([2, 2, 2, 2], [1, 1, -1, -1], [2, 2, 2], [1, 1, -1], [2, 3], [0, 1], None),
              ([2, 2, 2], [1, 1, -1], [2, 2, 2], [1, 1, -1], [2], [0], [3, 0, 2, 1]),
              ([3, 1, 2], [1, 1, -1], [2, 2, 2, 4], [1, -1, 1, -1], [2], [0], [4, 1, 0, 2, 3]),
              ([2, 2, 2, 2], [1, 1, -1, -1], [2, 2, 2, 2], [1, 1, -1, -1], [2, 3], [0, 1], [1, 0, 3, 2]),
              ([4, 2, 2, 4], [1, 1, -1, -1], [4, 2, 2], [-1, 1, -1], [0, 1], [0, 2], None)]
    for tja, da, tjb, db, ia, ib, perm in xcases:
        Xo = OM.xom(tja, da, tjb, db, ia, ib, perm)
        Xp = Q.su2_xom(tja, da, tjb, db, ia, ib, perm)
        assert Xo.shape == Xp.shape
        assert np.abs(Xo - Xp).max() < 1e-13
    # rank <= 3 everywhere: the scalar x3 (times the output permutation's sign)
    x = Q.su2_xom([2, 1, 3], [1, 1, -1], [3, 1, 2], [1, -1, -1], [2, 1], [0, 1])
    assert x.shape == (1, 1, 1) and abs(x[0, 0, 0] - O.x3([2, 1, 3], [1, 1, -1], [3, 1, 2], [1, -1, -1], [2, 1],
                                                           [0, 1])) < 1e-14


def _legs(rng, labels, dmin, dmax):
    return [(t, int(rng.randint(dmin, dmax + 1))) for t in labels]


def test_oracle_contract_om_full_space():
    """rank-3 x rank-3 over one leg (rank-4 result with outer multiplicity),
    rank-4 x rank-3, rank-4 x rank-4 over two legs, an output permutation and
    a conjugated operand, each against the full-space tensordot"""
    rng = np.random.RandomState(6)
    L = _legs(rng, [0, 1, 2, 3], 1, 3)
    S = [(1, 1)]
    A = gen.su2_om_tensor(11, [L, S, L], ["a", "s", "b*"])
    B = gen.su2_om_tensor(12, [L, S, L], ["b", "t", "c*"])
    C = OM.contract(A, [2], B, [0])
    assert C.rank == 4 and max(C.om()) > 1
    assert check_contraction(A, [2], B, [0], C) < 1e-13
    W = gen.su2_om_tensor(13, [[(0, 1), (2, 2)], [(0, 1), (2, 2)], S, S], ["w", "v*", "s*", "u*"])
    assert max(W.om()) > 1
    D = OM.contract(C, [1], W, [2])
    assert check_contraction(C, [1], W, [2], D) < 1e-13
    E = OM.contract(C, [1, 2], W, [2, 3], skip_zero=0)  # rank-4 x rank-4 over (s, t)
    assert check_contraction(C, [1, 2], W, [2, 3], E) < 1e-13
    P = OM.contract(A, [2], B, [0], perm=[3, 0, 2, 1])
    full = OM.contract(A, [2], B, [0])
    from fullspace import _layout, _sectors, expand
    lay = [_layout(_sectors(full, k, 0)) for k in range(4)]
    ref = np.transpose(expand(full, lay), [3, 0, 2, 1])
    lay2 = [lay[p] for p in [3, 0, 2, 1]]
    assert np.abs(expand(P, lay2) - ref).max() < 1e-13 * max(1, np.abs(ref).max())
    # conj(A) . A over (a, s): rank-2 x rank-2 of the kept legs
    G = OM.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert check_contraction(A, [0, 1], A, [0, 1], G) < 1e-13


def test_oracle_env_update_su2_full_space():
    """each contraction of the SU(2) environment update A'.(HK.W.A) (rank-4
    intermediates with outer multiplicity) against the full space"""
    A, HK, W = gen.su2_env_update(seed=21, tS=tuple(range(5)), dmin=1, dmax=3, dw=1)
    X1 = OM.contract(A, [0], HK, [0])
    assert check_contraction(A, [0], HK, [0], X1) < 1e-13
    X2 = OM.contract(X1, [1, 3], W, [2, 0])
    assert check_contraction(X1, [1, 3], W, [2, 0], X2) < 1e-13
    R = OM.contract(A, [0, 2], X2, [1, 3], conjA=1)
    assert R.rank == 3 and R.itags == ["R", "R*", "w2*"]
    assert check_contraction(A, [0, 2], X2, [1, 3], R) < 1e-13
