"""Outer multiplicity on the device (GPU tests): SU(2) contractions whose
operands or results have rank > 3 (contractDATA_group with tensorRefs,
Source/QSpace_aux.cc:27-345), against the oracle's block-wise restatement
(om.contract) and the full M-resolved space.

Bar: same label records, dims and outer multiplicities; data within 1e-12
relative Frobenius error.  The oracle's output blocks are in sorted record
order, so blocks are compared by record.
"""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from fullspace import check_contraction
from qspace_amd import gen

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _recs(T):
    return [tuple(tuple(int(x) for x in T.Q[k][i]) for k in range(T.rank)) for i in range(T.nblk)]


def assert_same_by_record(C, R):
    rc, rr = _recs(C), _recs(R)
    assert sorted(rc) == sorted(rr), (len(rc), len(rr))
    assert C.itags == R.itags
    dr = dict(zip(rr, R.data))
    num = den = 0.0
    for k, d in zip(rc, C.data):
        assert d.shape == dr[k].shape, (k, d.shape, dr[k].shape)
        num += float(np.sum(np.abs(d - dr[k]) ** 2))
        den += float(np.sum(np.abs(dr[k]) ** 2))
    assert np.sqrt(num) <= TOL * max(np.sqrt(den), 1e-300)
    wc = C.info.get("cgw")
    wr = R.info.get("cgw")
    if wc is not None and wr is not None:
        dwr = dict(zip(rr, wr[:, 0]))
        for k, w in zip(rc, wc[:, 0]):
            assert w == dwr[k]


def _legs(seed, labels, dmin, dmax):
    rng = np.random.RandomState(seed)
    return [(t, int(rng.randint(dmin, dmax + 1))) for t in labels]


def test_rank3_pair_to_rank4_result(gpu):
    """A(a,s,b*).B(b,t,c*): a rank-4 result whose blocks carry outer
    multiplicity, batched and plain engines, against the oracle and the full
    space"""
    L = _legs(1, range(0, 7), 2, 6)
    S = [(1, 1), (2, 1)]
    A = gen.su2_om_tensor(11, [L, S, L], ["a", "s", "b*"])
    B = gen.su2_om_tensor(12, [L, S, L], ["b", "t", "c*"])
    R = O.contract(A, [2], B, [0])
    assert max(R.om()) > 1
    for mode in ("batched", "plain"):
        C = Q.api._contract_pair(A, [2], 0, B, [0], 0, None, mode=mode)
        print(f"rank-4 result ({mode}): {C.nblk} blocks, max om {max(C.om())}", flush=True)
        assert_same_by_record(C, R)
    assert check_contraction(A, [2], B, [0], C) < 1e-13


def test_rank4_operands_perm_and_conj(gpu):
    """rank-4 x rank-4 over two legs, rank-4 x rank-3, an output permutation,
    and a conjugated rank-4 operand contracted to a rank-4 result"""
    L = _legs(2, range(0, 6), 2, 5)
    S = [(1, 1)]
    A = gen.su2_om_tensor(21, [L, S, S, L], ["a", "s", "t*", "b*"])
    wsec = [(0, 2), (2, 3)]
    W = gen.su2_om_tensor(22, [wsec, wsec, S, S], ["w", "v*", "t", "u*"])
    W2 = gen.su2_om_tensor(23, [wsec, wsec, S, S], ["w", "v*", "s*", "t"])
    W3 = gen.su2_om_tensor(24, [wsec, wsec, S, S], ["w", "v*", "s", "t*"])
    assert max(A.om()) > 1 and max(W.om()) > 1
    cases = [(A, [2], W, [2], None, 0, 0),
             (A, [1, 2], W2, [2, 3], None, 0, 0),
             (A, [1, 2], W3, [2, 3], [3, 1, 0, 2], 1, 0),
             (A, [3], A, [3], [2, 0, 5, 1, 4, 3], 0, 1),
             (A, [0, 1, 2], A, [0, 1, 2], None, 1, 0)]
    for X, ia, Y, ib, perm, cA, cB in cases:
        C = Q.api._contract_pair(X, ia, cA, Y, ib, cB, perm)
        R = O.contract(X, ia, Y, ib, perm=perm, conjA=cA, conjB=cB)
        print(f"case {ia}x{ib} perm={perm} conj=({cA},{cB}): {C.nblk} blocks", flush=True)
        assert_same_by_record(C, R)


def test_env_update_su2_cell_form(gpu):
    """SU(2) environment update L' = A'.(HK.W.A) in the cell form of
    updateHK.m ({A,'!2*',{{A,HK},W}}) on device-resident operands: rank-4
    intermediates with outer multiplicity and a rank-4 MPO, against the
    oracle's explicit chain and (small case) the full space"""
    for seed, tS, dmin, dmax, dw in ((31, tuple(range(5)), 1, 3, 1), (32, tuple(range(17)), 8, 64, 4)):
        A, HK, W = gen.su2_env_update(seed=seed, tS=tS, dmin=dmin, dmax=dmax, dw=dw)
        C = Q.contractQS([A, "!2*", [[A, HK], W]])
        X1 = O.contract(A, [0], HK, [0])            # (R*, s, L, w*)
        X2 = O.contract(X1, [1, 3], W, [2, 0])      # (R*, L, w2*, s)
        R = O.contract(A, [0, 2], X2, [1, 3], conjA=1)
        print(f"env update SU(2) 2S<={max(tS)}: {C.nblk} blocks", flush=True)
        assert C.itags == ["R", "R*", "w2*"]
        assert_same_by_record(C, R)
        if seed == 31:
            assert check_contraction(A, [0, 2], X2, [1, 3], C) < 1e-13


def test_om_plan_reuse_and_refusals(gpu):
    """a plan with outer multiplicity executes twice on new data; permuteQS
    and plusQS refuse blocks with outer multiplicity"""
    L = _legs(3, range(0, 5), 2, 4)
    S = [(1, 1)]
    A = gen.su2_om_tensor(31, [L, S, L], ["a", "s", "b*"])
    B = gen.su2_om_tensor(32, [L, S, L], ["b", "t", "c*"])
    P = Q.CtrPlan(A, "3", B, "1")
    dA = Q.DeviceBuffer.from_array(A.packed())
    for seed in (1, 2):
        B2 = B.with_packed(np.random.RandomState(seed).standard_normal(B.numel()))
        dB = Q.DeviceBuffer.from_array(B2.packed())
        dC = Q.DeviceBuffer(P.out.numel())
        P.execute(dA, dB, dC)
        got = P.out.with_packed(dC.download())
        R = O.contract(A, [2], B2, [0], skip_zero=0)
        assert_same_by_record(got, R)
    C = Q.contractQS(A, "3", B, "1")
    with pytest.raises(Q.QsxError, match="outer multiplicity"):
        Q.permuteQS(C, [2, 1, 3, 4])
    with pytest.raises(Q.QsxError, match="outer multiplicity"):
        Q.plusQS(C, C)
