"""CG indexing and block structure of the device label engine against the
oracle, at the workloads' full sizes (GPU tests).

contract_getIdxSet (Source/QSpace.cc:4131-4294) defines the pairs (ia, ib)
of every output block, their order (matchIndex A-outer / B-inner,
Source/wbsort.cc:262-263; groupRecs, Source/wbMatrix.cc:1682-1731) and the
output label records; a plan's index set (qsx_ctr_plan_idxset) must equal
the oracle's bit for bit.  The per-pair factor afac = X-symbols x cgw comes
from two independent constructions (the product's exact Racah evaluation,
the oracle's lowering-operator CGTs), so it is compared to 1e-13 relative,
with the same exact zeros.
"""
import numpy as np
import pytest

import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from helpers import assert_same_structure, rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _compare(A, ica, B, icb, conjA=0, conjB=0):
    ca = "".join(str(i + 1) for i in ica) + ("*" if conjA else "")
    cb = "".join(str(i + 1) for i in icb) + ("*" if conjB else "")
    P = Q.CtrPlan(A, ca, B, cb)
    ia, ib, af, seg, qc = P.idxset()
    ra, rb, raf, rseg, rqc = O.pairs(A, ica, B, icb, conjA=conjA, conjB=conjB)
    print(f"idxset {ca} x {cb}: {len(ia)} pairs, {len(seg) - 1} output blocks", flush=True)
    assert len(ia) == len(ra) and len(seg) == len(rseg)
    assert np.array_equal(ia, ra)
    assert np.array_equal(ib, rb)
    assert np.array_equal(seg, rseg)
    assert np.array_equal(qc, rqc)
    assert np.array_equal(af == 0, raf == 0)
    nz = raf != 0
    assert np.all(np.abs(af[nz] - raf[nz]) <= 1e-13 * np.abs(raf[nz]))
    return P


def test_idxset_u1_overlap(gpu):
    """config 1 at full size: contractQS(A,'12*',A,'12')"""
    A = gen.u1_mps()
    _compare(A, [0, 1], A, [0, 1], conjA=1)


def test_idxset_and_data_su2_env_full_size(gpu):
    """config 2 at full size (2S in 0..16, dims 8..64, ~600 multiplets per
    bond): index sets of L.A and X.conj(A), and both results' data within
    1e-12 of the oracle"""
    A, L = gen.su2_mps()
    assert sum(int(d.shape[0]) for d in L.data) >= 590
    P1 = _compare(L, [1], A, [0])
    X, XR = Q.contractQS(L, "2", A, "1"), O.contract(L, [1], A, [0])
    assert_same_structure(X, XR)
    assert rel_err(X, XR) < TOL
    _compare(XR, [0, 1], A, [0, 1], conjB=1)
    C, CR = Q.contractQS(X, "12", A, "12*"), O.contract(XR, [0, 1], A, [0, 1], conjB=1)
    assert_same_structure(C, CR)
    assert rel_err(C, CR) < TOL
    assert P1.flops > 0


@pytest.mark.parametrize("kind", ["heis4000", "hubbard4000"])
def test_idxset_hpsi_stages(gpu, kind):
    """both H*psi stages of the headline operators (main and Hermitian-
    conjugate term): HL.psi over (E1l*, E1l), conj(HL).psi over leg 1, and
    T.HR over (w, E2r)"""
    if kind == "heis4000":
        HL, psi, HR = gen.heis_hpsi(seed=1003, Dstar=4000)
    else:
        HL, psi, HR = gen.hubbard_hpsi(seed=1005, Dstar=4000)
    P1 = _compare(HL, [1], psi, [0])
    _compare(HL, [0], psi, [0], conjA=1)
    T = P1.out
    _compare(T, [1, 2], HR, [2, 1])


def test_hpsi_heis4000_matches_oracle(gpu, heis4000):
    """the headline H*psi (Heisenberg, D* = 4000, both terms) under the
    autotuned policy against the oracle's explicit contraction chain"""
    HL, psi, HR, x, ref = heis4000
    H = Q.HPsi(HL, psi, HR, hconj=1)
    H.autotune(reps=3)
    got = H.apply(x)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < TOL
