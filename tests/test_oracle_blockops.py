"""Pins of the oracle's permuteQS / plusQS restatements (CPU only).

  * permute: Full(permute(A, p)) = transpose(Full(A), p) in the full
    M-resolved space (convention-independent), and a conj permutation
    contracts exactly like the contraction's own conj flag;
  * plus: Full(A + bfac B) = Full(A) + bfac Full(B), with weights other than
    the external norm included;
  * block order and zero handling (the reference's plus_plain / Cat +
    MakeUnique / SkipZeroData rules) by hand-built cases whose expected
    order is written out literally.
"""
import numpy as np
import pytest

import oracle as O
from fullspace import _layout, _sectors, _syms, expand
from qspace_amd import QSpace, gen


def u1_rank4(seed, nq=3, dmin=1, dmax=6):
    """U(1) A(a,b,c,d) with qa + qb = qc + qd."""
    rng = np.random.RandomState(seed)
    qs = list(range(-nq, nq + 1))
    dim = [{q: int(rng.randint(dmin, dmax + 1)) for q in qs} for _ in range(4)]
    Q, data = [[], [], [], []], []
    for a in qs:
        for b in qs:
            for c in qs:
                d = a + b - c
                if d not in dim[3] or rng.rand() < 0.3:
                    continue
                for k, v in enumerate((a, b, c, d)):
                    Q[k].append([v])
                shp = (dim[0][a], dim[1][b], dim[2][c], dim[3][d])
                data.append(rng.standard_normal(shp))
    return QSpace([np.array(q) for q in Q], data, "A", ["a", "b", "c*", "d*"])


def su2_env(seed, Dstar=60):
    HL, psi, HR = gen.heis_hpsi(seed=seed, Dstar=Dstar, mu=1.6, sigma=0.5, dmax=12)
    return HL


def _full(T, layouts):
    return expand(T, layouts)


def _layouts(T):
    j = next((i for i, s in enumerate(_syms(T)) if s == "SU2"), None)
    return [_layout(_sectors(T, k, j)) for k in range(T.rank)]


def _union_layouts(A, B):
    j = next((i for i, s in enumerate(_syms(A)) if s == "SU2"), None)
    out = []
    for k in range(A.rank):
        u = dict(_sectors(A, k, j))
        u.update(_sectors(B, k, j))
        out.append(_layout(u))
    return out


@pytest.mark.parametrize("perm", [[1, 0, 2, 3], [3, 2, 1, 0], [2, 0], [0, 1, 3, 2], [1, 2, 3, 0]])
def test_permute_u1_fullspace(perm):
    A = u1_rank4(3)
    P = O.permute(A, perm)
    p = list(perm) + [i for i in range(4) if i not in perm]
    lay = _layouts(A)
    FA = _full(A, lay)
    FP = _full(P, [lay[k] for k in p])
    assert np.array_equal(FP, np.transpose(FA, p))
    assert P.itags == [A.itags[k] for k in p]


@pytest.mark.parametrize("perm", [[1, 0, 2], [2, 1, 0], [1, 2, 0], [0, 2, 1]])
def test_permute_su2_fullspace(perm):
    A = su2_env(7)
    P = O.permute(A, perm)
    lay = _layouts(A)
    FA = _full(A, lay)
    FP = _full(P, [lay[k] for k in perm])
    assert np.abs(FP - np.transpose(FA, perm)).max() < 1e-13 * max(1.0, np.abs(FA).max())


def test_permute_normcgw_factor():
    """weights given as 3 x normExt: NormCGW moves the factor 3 into the data."""
    A = su2_env(8)
    dirs = [(-1 if t.endswith("*") else 1) for t in A.itags]
    cg = np.array([[3.0 * O.norm_ext([int(A.Q[k][i, 0]) for k in range(3)], dirs)] for i in range(A.nblk)])
    A3 = QSpace(A.Q, A.data, A.qtype, A.itags, cg)
    P = O.permute(A3, [2, 0, 1])
    P1 = O.permute(A, [2, 0, 1])
    for a, b in zip(P.data, P1.data):
        assert np.allclose(a, 3.0 * b, rtol=1e-15, atol=0)
    assert np.array_equal(P.info["cgw"], P1.info["cgw"])
    lay = _layouts(A)
    assert np.abs(_full(P, [lay[k] for k in (2, 0, 1)]) -
                  np.transpose(_full(A3, lay), (2, 0, 1))).max() < 1e-12


def test_permute_identity_and_errors():
    A = u1_rank4(4)
    P = O.permute(A, [])
    assert all(np.array_equal(a, b) for a, b in zip(P.data, A.data))
    with pytest.raises(O.OracleError):
        O.permute(A, [0, 1, 2, 3, 0])
    with pytest.raises(O.OracleError):
        O.permute(A, [1, 1])


@pytest.mark.parametrize("kind", ["u1", "su2"])
def test_permute_conj_contracts_like_conj_flag(kind):
    """contract(permuteQS(A,[],'conj'), B) = contract(A*, B): the conj of a
    permutation and of a contraction agree."""
    if kind == "u1":
        A = gen.u1_mps(seed=21, Lq=tuple(range(-4, 5, 2)), Rq=tuple(range(-5, 6, 2)), dmin=2, dmax=6)
    else:
        A, _ = gen.su2_mps(seed=22, tS=tuple(range(6)), dmin=2, dmax=6)
    R1 = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    Ac = O.permute(A, [], conj=1)
    assert Ac.itags == [t[:-1] if t.endswith("*") else t + "*" for t in A.itags]
    R2 = O.contract(Ac, [0, 1], A, [0, 1])
    assert R1.itags == R2.itags
    assert np.array_equal(R1.QIDX, R2.QIDX)
    for a, b in zip(R1.data, R2.data):
        assert np.abs(a - b).max() <= 1e-14 * max(1.0, np.abs(a).max())


def _split(T, lo, hi, seed):
    """blocks [lo, hi) of T with fresh data"""
    rng = np.random.RandomState(seed)
    Q = [q[lo:hi] for q in T.Q]
    data = [rng.standard_normal(d.shape) for d in T.data[lo:hi]]
    return QSpace(Q, data, T.qtype, T.itags)


@pytest.mark.parametrize("kind,bfac", [("u1", 1.0), ("u1", -0.375), ("su2", 1.0), ("su2", 2.5)])
def test_plus_fullspace(kind, bfac):
    T = u1_rank4(5) if kind == "u1" else su2_env(9)
    n = T.nblk
    A = _split(T, 0, 2 * n // 3, 1)
    B = _split(T, n // 3, n, 2)
    C = O.plus(A, B, bfac)
    lay = _union_layouts(A, B)
    ref = _full(A, lay) + bfac * _full(B, lay)
    assert np.abs(_full(C, lay) - ref).max() <= 1e-14 * max(1.0, np.abs(ref).max())
    assert C.nblk == n


def test_plus_su2_weights():
    """B with weights 2 x normExt counts twice."""
    T = su2_env(10)
    n = T.nblk
    A = _split(T, 0, n, 1)
    B = _split(T, 0, n, 2)
    dirs = [(-1 if t.endswith("*") else 1) for t in B.itags]
    cg = np.array([[2.0 * O.norm_ext([int(B.Q[k][i, 0]) for k in range(3)], dirs)] for i in range(n)])
    B2 = QSpace(B.Q, B.data, B.qtype, B.itags, cg)
    C = O.plus(A, B2, 0.5)
    for c, a, b in zip(C.data, A.data, B.data):
        assert np.allclose(c, a + b, rtol=1e-15, atol=1e-300)


def _tiny(records, qtype, tags, vals, rank=3):
    Q = [np.array([[r[k]] for r in records]) for k in range(rank)]
    data = [np.full((1,) * rank, float(v)) for v in vals]
    return QSpace(Q, data, qtype, tags)


def test_plus_order_rules():
    # abelian: always the sorted union
    A = _tiny([(2, 0, 2), (0, 1, 1)], "A", ["a", "b", "c*"], [1, 2])
    B = _tiny([(1, 0, 1), (-1, 1, 0)], "A", ["a", "b", "c*"], [3, 4])
    C = O.plus(A, B)
    assert [tuple(r) for r in C.QIDX] == [(-1, 1, 0), (0, 1, 1), (1, 0, 1), (2, 0, 2)]
    assert [float(d.ravel()[0]) for d in C.data] == [4.0, 2.0, 3.0, 1.0]
    # SU(2), no common record: A's blocks then B's, each in its own order
    A = _tiny([(2, 1, 1), (0, 1, 1)], "SU2", ["a", "b", "c*"], [1, 2])
    B = _tiny([(1, 1, 2), (1, 1, 0)], "SU2", ["a", "b", "c*"], [3, 4])
    C = O.plus(A, B)
    assert [tuple(r) for r in C.QIDX] == [(2, 1, 1), (0, 1, 1), (1, 1, 2), (1, 1, 0)]
    # SU(2) with a common record: sorted unique records
    B = _tiny([(1, 1, 2), (0, 1, 1)], "SU2", ["a", "b", "c*"], [3, 4])
    C = O.plus(A, B)
    assert [tuple(r) for r in C.QIDX] == [(0, 1, 1), (1, 1, 2), (2, 1, 1)]
    assert [float(d.ravel()[0]) for d in C.data] == [6.0, 3.0, 1.0]


def test_plus_zero_blocks():
    A = _tiny([(0, 1, 1), (1, 0, 1)], "A", ["a", "b", "c*"], [1, 2])
    B = _tiny([(0, 1, 1)], "A", ["a", "b", "c*"], [1])
    C = O.plus(A, B, -1.0)           # (0,1,1) cancels exactly: dropped (rank 3)
    assert [tuple(r) for r in C.QIDX] == [(1, 0, 1)]
    C = O.plus(A, B, -1.0, skip_zero=0)
    assert C.nblk == 2
    # rank-2 block-diagonal matrix keeps its zero blocks
    M = _tiny([(0, 0), (1, 1)], "A", ["a", "a*"], [1, 2], rank=2)
    N = _tiny([(0, 0)], "A", ["a", "a*"], [1], rank=2)
    assert O.plus(M, N, -1.0).nblk == 2
    # SU(2): a zero block of a group is not added (the non-zero one is kept)
    A = _tiny([(0, 1, 1), (2, 1, 1)], "SU2", ["a", "b", "c*"], [0, 2])
    B = _tiny([(0, 1, 1)], "SU2", ["a", "b", "c*"], [5])
    C = O.plus(A, B, 1.0)
    assert float(C.data[0].ravel()[0]) == 5.0


def test_plus_errors_and_tags():
    A = _tiny([(0, 1, 1)], "A", ["a", "b", "c*"], [1])
    B = _tiny([(0, 1, 1)], "A", ["a", "b*", "c*"], [1])
    with pytest.raises(O.OracleError):
        O.plus(A, B)
    B = _tiny([(0, 1, 1)], "A", ["x", "b", "c*"], [1])
    assert O.plus(A, B).itags == ["", "", "*"]       # names cleared, conj kept
    S = _tiny([(0, 1, 1)], "SU2", ["a", "b", "c*"], [1])
    S2 = _tiny([(0, 1, 1)], "SU2", ["x", "b", "c*"], [1])
    with pytest.raises(O.OracleError):
        O.plus(S, S2)
    Bd = QSpace([np.array([[0]]), np.array([[1]]), np.array([[1]])], [np.ones((2, 1, 1))], "A", ["a", "b", "c*"])
    with pytest.raises(O.OracleError):
        O.plus(A, Bd)


@pytest.mark.parametrize("perm", [[1, 0, 2], [2, 1, 0], [0, 2, 1], [2, 0, 1]])
def test_contract_output_perm_su2_fullspace(perm):
    """contractQS(A, ia, B, ib, perm) with SU(2) labels: Full(C) equals the
    permuted tensordot of the full operands, also when the permutation
    reorders legs of equal label and direction (the weight takes the sign)."""
    HL, psi, HR = gen.heis_hpsi(seed=7, Dstar=60, mu=1.6, sigma=0.5, dmax=12)
    A, B, ia, ib = HL, psi, [1], [0]
    lay = _union_layouts_pair(A, ia, B, ib)
    ref = np.tensordot(_full(A, lay[0]), _full(B, lay[1]), axes=(ia, ib))
    layC = [lay[0][k] for k in range(A.rank) if k not in ia] + [lay[1][k] for k in range(B.rank) if k not in ib]
    C = O.contract(A, ia, B, ib, perm=perm)
    FC = _full(C, [layC[k] for k in perm])
    assert np.abs(FC - np.transpose(ref, perm)).max() <= 1e-13 * np.abs(ref).max()


def _union_layouts_pair(A, ia, B, ib):
    j = next((i for i, s in enumerate(_syms(A)) if s == "SU2"), None)
    LA = [_sectors(A, k, j) for k in range(A.rank)]
    LB = [_sectors(B, k, j) for k in range(B.rank)]
    for a, b in zip(ia, ib):
        u = dict(LA[a])
        u.update(LB[b])
        LA[a] = u
        LB[b] = u
    return [_layout(s) for s in LA], [_layout(s) for s in LB]


def _cplx(T, seed):
    rng = np.random.RandomState(seed)
    return QSpace(T.Q, [d + 1j * rng.standard_normal(d.shape) for d in T.data], T.qtype, T.itags,
                  T.info.get("cgw"))


def test_complex_permute_and_plus_full_space():
    """complex U(1) data: Full(permute(A,p)) = transpose(Full(A)), with conj
    the complex conjugate; Full(A + b B) = Full(A) + b Full(B)"""
    A = _cplx(u1_rank4(21), 1)
    lay = [_layout(_sectors(A, k, None)) for k in range(A.rank)]
    F = expand(A, lay)
    for perm, conj in (([2, 0, 3, 1], 0), ([1, 0, 2, 3], 1), ([0, 1, 2, 3], 1)):
        P = O.permute(A, perm, conj=conj)
        ref = np.transpose(F, perm)
        if conj:
            ref = np.conj(ref)
        assert np.abs(expand(P, [lay[k] for k in perm]) - ref).max() < 1e-14
    B = QSpace(A.Q, [np.asarray(d).real * 0.5 for d in A.data], A.qtype, A.itags)
    C = O.plus(A, B, -2.0)
    assert np.abs(expand(C, lay) - (F - 2.0 * expand(B, lay))).max() < 1e-13
