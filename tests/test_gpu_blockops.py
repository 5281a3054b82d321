"""permuteQS / plusQS on the device vs the CPU oracle (GPU tests).

Bar: structure (Q records, block order, dims, itags, cgw) and data
bit-exact -- both operations are data movement plus at most one rounded
product and one rounded sum per element, evaluated in the same order.
"""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import QSpace, gen
from helpers import assert_same_structure
from test_oracle_blockops import _split, _tiny, su2_env, u1_rank4

pytestmark = pytest.mark.gpu


def _exact(C, R):
    assert_same_structure(C, R)
    for i, (a, b) in enumerate(zip(C.data, R.data)):
        assert a.shape == b.shape, i
        assert np.array_equal(a, b), (i, float(np.abs(a - b).max()))


CASES = [
    ("u1_small", [1, 0, 2, 3], 0), ("u1_small", [3, 2, 1, 0], 0), ("u1_small", [2, 0], 1),
    ("u1_big", [1, 0, 2, 3], 0), ("u1_big", [0, 2, 1, 3], 0), ("u1_big", [3, 1, 2, 0], 0), ("u1_big", [1, 2, 3, 0], 1),
    ("su2", [2, 1, 0], 0), ("su2", [1, 2, 0], 0), ("su2", [0, 2, 1], 1), ("su2", [], 1),
]


def _tensor(kind):
    if kind == "u1_small":
        return u1_rank4(3)
    if kind == "u1_big":
        return u1_rank4(4, nq=2, dmin=6, dmax=20)
    return su2_env(7, Dstar=300)


@pytest.mark.parametrize("kind,perm,conj", CASES)
def test_permute_matches_oracle(gpu, kind, perm, conj):
    A = _tensor(kind)
    print(f"permute {kind} perm={perm} conj={conj} blocks={A.nblk}", flush=True)
    C = Q.permuteQS(A, perm, conj=bool(conj))
    R = O.permute(A, perm, conj=conj)
    _exact(C, R)


def test_permute_round_trip_full_size(gpu):
    """D* = 4000 Heisenberg environment: permute then the inverse permutation
    gives the input back bit for bit (size-independent property)."""
    HL, psi, HR = gen.heis_hpsi(seed=1003, Dstar=4000)
    P = Q.permuteQS(HL, [2, 0, 1])
    B = Q.permuteQS(P, [1, 2, 0])
    assert np.array_equal(B.QIDX, HL.QIDX)
    for a, b in zip(B.data, HL.data):
        assert np.array_equal(a, b)
    R = O.permute(HL, [2, 0, 1])
    _exact(P, R)


PLUS = [("u1", 1.0), ("u1", -0.375), ("su2", 1.0), ("su2", 2.5)]


@pytest.mark.parametrize("kind,bfac", PLUS)
def test_plus_matches_oracle(gpu, kind, bfac):
    T = u1_rank4(5, dmin=2, dmax=12) if kind == "u1" else su2_env(9, Dstar=300)
    n = T.nblk
    A = _split(T, 0, 2 * n // 3, 1)
    B = _split(T, n // 3, n, 2)
    print(f"plus {kind} bfac={bfac} blocks={A.nblk}+{B.nblk}", flush=True)
    _exact(Q.plusQS(A, B, bfac), O.plus(A, B, bfac))


def test_plus_rules_match_oracle(gpu):
    A = _tiny([(2, 1, 1), (0, 1, 1)], "SU2", ["a", "b", "c*"], [1, 2])
    for B in (_tiny([(1, 1, 2), (1, 1, 0)], "SU2", ["a", "b", "c*"], [3, 4]),
              _tiny([(1, 1, 2), (0, 1, 1)], "SU2", ["a", "b", "c*"], [3, -2])):
        _exact(Q.plusQS(A, B), O.plus(A, B))
        _exact(Q.plusQS(A, B, skip_zero=0), O.plus(A, B, skip_zero=0))
    A = _tiny([(0, 1, 1), (1, 0, 1)], "A", ["a", "b", "c*"], [1, 2])
    B = _tiny([(0, 1, 1)], "A", ["x", "b", "c*"], [1])
    _exact(Q.plusQS(A, B, -1.0), O.plus(A, B, -1.0))
    E = QSpace([np.zeros((0, 1), np.int32)] * 3, [], "A", ["a", "b", "c*"])
    _exact(Q.plusQS(E, B, 2.0), O.plus(E, B, 2.0))
    with pytest.raises(Q.QsxError):
        Q.plusQS(A, _tiny([(0, 1, 1)], "A", ["a", "b*", "c*"], [1]))


def test_plus_weights_match_oracle(gpu):
    T = su2_env(10, Dstar=200)
    n = T.nblk
    A = _split(T, 0, n, 1)
    B = _split(T, 0, n, 2)
    dirs = [(-1 if t.endswith("*") else 1) for t in B.itags]
    cg = np.array([[2.0 * O.norm_ext([int(B.Q[k][i, 0]) for k in range(3)], dirs)] for i in range(n)])
    B2 = QSpace(B.Q, B.data, B.qtype, B.itags, cg)
    _exact(Q.plusQS(A, B2, 0.5), O.plus(A, B2, 0.5))


@pytest.mark.parametrize("perm", [[2, 1, 0], [1, 0, 2], [2, 0, 1]])
def test_contract_output_perm_su2(gpu, perm):
    """contractQS with a user permutation of an SU(2) result: weights take the
    sign of the CGT of the new leg order (device vs oracle)."""
    HL, psi, HR = gen.heis_hpsi(seed=7, Dstar=200, mu=2.0, sigma=0.6, dmax=24)
    C = Q.api._contract_pair(HL, [1], 0, psi, [0], 0, perm)
    R = O.contract(HL, [1], psi, [0], perm=perm)
    assert_same_structure(C, R)
    num = sum(float(np.sum((a - b) ** 2)) for a, b in zip(C.data, R.data))
    den = sum(float(np.sum(b * b)) for b in R.data)
    assert np.sqrt(num / den) < 1e-12
    # the case with a sign occurs: some blocks are minus the transposed
    # blocks of the unpermuted result
    U = O.contract(HL, [1], psi, [0])
    flips = sum(np.array_equal(c, -np.transpose(u, perm)) and np.any(u != 0) for c, u in zip(R.data, U.data))
    assert flips > 0


def test_cell_form_perm_on_device(gpu):
    """contractQS({HL, psi}, HR, perm): itag contractions, then the device
    permute of the result."""
    HL, psi, HR = gen.heis_hpsi(seed=8, Dstar=150, mu=2.0, sigma=0.6, dmax=24)
    C0 = Q.contractQS([HL, psi], HR)
    C = Q.contractQS([HL, psi], HR, [2, 1])
    _exact(C, O.permute(C0, [1, 0]))


def _cplx(T, seed):
    rng = np.random.RandomState(seed)
    return QSpace(T.Q, [np.asfortranarray(d + 1j * rng.standard_normal(d.shape)) for d in T.data], T.qtype,
                  T.itags, T.info.get("cgw"))


@pytest.mark.parametrize("kind,perm,conj", [("u1_small", [1, 0, 2, 3], 0), ("u1_big", [3, 1, 2, 0], 1),
                                             ("u1_small", [], 1), ("su2", [2, 1, 0], 0), ("su2", [0, 2, 1], 1)])
def test_complex_permute_matches_oracle(gpu, kind, perm, conj):
    """both planes through the same permutation; conj conjugates the data"""
    A = _cplx(_tensor(kind), 5)
    C = Q.permuteQS(A, perm, conj=bool(conj))
    assert C.is_complex()
    _exact(C, O.permute(A, perm, conj=conj))


@pytest.mark.parametrize("kind,which", [("u1", "both"), ("u1", "B"), ("su2", "A"), ("su2", "both")])
def test_complex_plus_matches_oracle(gpu, kind, which):
    T = u1_rank4(5, dmin=2, dmax=12) if kind == "u1" else su2_env(9, Dstar=300)
    n = T.nblk
    A = _split(T, 0, 2 * n // 3, 1)
    B = _split(T, n // 3, n, 2)
    if which in ("both", "A"):
        A = _cplx(A, 3)
    if which in ("both", "B"):
        B = _cplx(B, 4)
    _exact(Q.plusQS(A, B, -0.75), O.plus(A, B, -0.75))
