"""Pins of the CPU oracle (CPU only).

Each check is independent of the oracle's own formulas:
  * SU(2) CG coefficients against sympy's exact Condon-Shortley values
    (tests/golden/su2_cg.json, made by tests/golden/make_golden.py) and
    against orthonormality at large spin;
  * CGT conventions against the restated rule (norm 1, first non-zero entry
    in column-major canonical order positive, canonical leg order);
  * X-symbols / cgw bookkeeping against the full M-resolved space;
  * the label engine against a brute-force matcher with the order stated
    independently;
  * U(1) data against a dense embedding + numpy.tensordot.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from fullspace import check_contraction
from qspace_amd import QSpace, gen

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_cg_against_sympy_fixture():
    with open(os.path.join(GOLDEN, "su2_cg.json")) as f:
        rows = json.load(f)["rows"]
    assert len(rows) > 300
    err = max(abs(O.su2_cg(*r[:6]) - r[6]) for r in rows)
    assert err < 1e-14


@pytest.mark.parametrize("tj1,tj2", [(1, 60), (2, 101), (3, 40), (8, 9)])
def test_cg_orthonormal_large_spin(tj1, tj2):
    """sum_{m1,m2} <j1 m1 j2 m2|J M><j1 m1 j2 m2|J' M> = delta_JJ' (size-independent)."""
    Js = list(range(abs(tj1 - tj2), tj1 + tj2 + 1, 2))
    for tM in (tj1 + tj2 - 2 * ((tj1 + tj2) // 3), (tj1 + tj2) % 2):
        rows = []
        for tJ in Js:
            if abs(tM) > tJ:
                continue
            v = [O.su2_cg(tj1, tm1, tj2, tM - tm1, tJ, tM) for tm1 in range(-tj1, tj1 + 1, 2)]
            rows.append(v)
        G = np.array(rows) @ np.array(rows).T
        assert np.abs(G - np.eye(len(rows))).max() < 1e-13


def _first_nonzero_colmajor(T):
    v = T.ravel(order="F")
    nz = np.nonzero(np.abs(v) > 1e-14)[0]
    return v[nz[0]]


@pytest.mark.parametrize("tj,dirs", [
    ([1, 1, 0], [1, 1, -1]), ([2, 3, 5], [1, 1, -1]), ([4, 2, 4], [1, -1, 1]), ([3, 3], [1, -1]),
    ([3, 3], [1, 1]), ([2, 2], [-1, -1]), ([5, 1, 4], [-1, -1, 1]), ([2, 2, 2], [1, 1, 1])])
def test_block_cgt_convention(tj, dirs):
    """|C|_F = 1; first non-zero in canonical column-major order positive;
    the block tensor is the canonical one with legs permuted back."""
    T = O.block_cgt(tj, dirs)
    assert T.shape == tuple(t + 1 for t in tj)
    assert abs(np.linalg.norm(T) - 1) < 1e-14
    r = len(tj)
    nin = sum(d > 0 for d in dirs)
    flip = nin < (r + 1) // 2
    flag = [((0 if d > 0 else 1) ^ (1 if flip else 0)) for d in dirs]
    order = sorted(range(r), key=lambda k: (flag[k], tj[k], k))
    canon = np.transpose(T, order)
    assert _first_nonzero_colmajor(canon) > 0


def test_block_cgt_is_cg():
    """rank-3 (in,in,out) CGT = CG coefficient / sqrt(2j3+1) up to the sign rule."""
    tj = [2, 3, 3]
    T = O.block_cgt(tj, [1, 1, -1])
    ref = np.zeros(T.shape)
    for i1 in range(3):
        for i2 in range(4):
            for i3 in range(4):
                ref[i1, i2, i3] = O.su2_cg(2, 2 - 2 * i1, 3, 3 - 2 * i2, 3, 3 - 2 * i3)
    ref /= 2.0
    s = np.sign(_first_nonzero_colmajor(ref))
    assert np.abs(T - s * ref).max() < 1e-14


def test_norm_ext():
    assert O.norm_ext([3, 3], [1, -1]) == np.sqrt(4)
    assert O.norm_ext([2, 1, 3], [1, 1, -1]) == np.sqrt(4)
    assert O.norm_ext([2, 1, 3], [-1, 1, 1]) == np.sqrt(3)
    assert O.norm_ext([2, 2, 2], [1, 1, 1]) == 1.0


# ---------------------------------------------------------------- full space
def test_fullspace_su2_env():
    A, L = gen.su2_mps(seed=31, tS=tuple(range(5)), dmin=2, dmax=4)
    X = O.contract(L, [1], A, [0])
    assert check_contraction(L, [1], A, [0], X) < 1e-13
    C = O.contract(X, [0, 1], A, [0, 1], conjB=1)
    assert check_contraction(X, [0, 1], A, [0, 1], C) < 1e-13


def test_fullspace_su2_hpsi_terms():
    HL, psi, HR = gen.heis_hpsi(seed=32, Dstar=14, mu=1.0, sigma=0.4, dmax=4)
    T = O.contract(HL, [1], psi, [0], skip_zero=0)
    assert check_contraction(HL, [1], psi, [0], T) < 1e-13
    H = O.contract(T, [1, 2], HR, [2, 1], skip_zero=0)
    assert check_contraction(T, [1, 2], HR, [2, 1], H) < 1e-13
    T2 = O.contract(HL, [0], psi, [0], conjA=1, skip_zero=0)
    assert check_contraction(HL, [0], psi, [0], T2) < 1e-13
    H2 = O.contract(T2, [1, 2], HR, [2, 0], conjB=1, skip_zero=0)
    assert check_contraction(T2, [1, 2], HR, [2, 0], H2) < 1e-13


def test_fullspace_pseudo_mpo_terms():
    """every contraction of the pseudo-MPO H*psi (get_HPsi.m:62-83: HK_L.psi,
    psi.HK_R, {OP_L,'*',psi}.OP_R and {OP_L,psi}.OP_R') equals its full
    M-space tensordot, and their block-wise sum equals the sum of the full
    tensors"""
    from helpers import pseudo_hpsi_oracle
    from fullspace import _layout, _sectors, expand
    HKL, psi, HKR, OPL, OPR = gen.heis_pseudo_hpsi(seed=34, Dstar=14, mu=1.0, sigma=0.4, dmax=4, nop=1)
    L, R = OPL[0], OPR[0]
    assert check_contraction(HKL, [1], psi, [0], O.contract(HKL, [1], psi, [0], skip_zero=0)) < 1e-13
    assert check_contraction(psi, [1], HKR, [1], O.contract(psi, [1], HKR, [1], skip_zero=0)) < 1e-13
    T1 = O.contract(L, [0], psi, [0], conjA=1, skip_zero=0)
    assert check_contraction(L, [0], psi, [0], T1) < 1e-13
    X1 = O.contract(T1, [1, 2], R, [2, 1], skip_zero=0)
    assert check_contraction(T1, [1, 2], R, [2, 1], X1) < 1e-13
    T2 = O.contract(L, [1], psi, [0], skip_zero=0)
    assert check_contraction(L, [1], psi, [0], T2) < 1e-13
    X2 = O.contract(T2, [1, 2], R, [2, 0], conjB=1, skip_zero=0)
    assert check_contraction(T2, [1, 2], R, [2, 0], X2) < 1e-13
    # the sum, in the full space of psi's legs
    H = pseudo_hpsi_oracle(HKL, psi, HKR, OPL, OPR, [1])
    lay = [_layout(_sectors(psi, k, 0)) for k in range(2)]
    full = sum(expand(T, lay) for T in (O.contract(HKL, [1], psi, [0], skip_zero=0),
                                        O.contract(psi, [1], HKR, [1], skip_zero=0), X1, X2))
    assert np.abs(expand(H, lay) - full).max() < 1e-13 * max(1.0, np.abs(full).max())


def test_fullspace_su2_u1_hubbard_terms():
    HL, psi, HR = gen.hubbard_hpsi(seed=33, Dstar=14, mu=1.0, sigma=0.4, dmax=3, qmax=2)
    T = O.contract(HL, [1], psi, [0], skip_zero=0)
    assert check_contraction(HL, [1], psi, [0], T) < 1e-13
    H = O.contract(T, [1, 2], HR, [2, 1], skip_zero=0)
    assert check_contraction(T, [1, 2], HR, [2, 1], H) < 1e-13


def test_fullspace_rank2_chain_and_scalar():
    rng = np.random.RandomState(3)
    labels = [0, 1, 2, 3]
    X = QSpace([[[q] for q in labels], [[q] for q in labels]], [rng.standard_normal((2, 3)) for _ in labels],
               "SU2", ["a", "b*"])
    Y = QSpace([[[q] for q in labels], [[q] for q in labels]], [rng.standard_normal((3, 2)) for _ in labels],
               "SU2", ["b", "c*"])
    Z = O.contract(X, [1], Y, [0])
    assert check_contraction(X, [1], Y, [0], Z) < 1e-13
    S = O.contract(X, [0, 1], X, [0, 1], conjA=1)  # full contraction to a scalar
    assert S.rank == 0 and len(S.data) == 1
    ref = sum((q + 1) * float(np.sum(d * d)) for q, d in zip(labels, X.data))
    assert abs(float(np.ravel(S.data[0])[0]) - ref) < 1e-12 * ref


# ---------------------------------------------------------------- U(1) dense
def _dense_u1(T, offs):
    full = np.zeros(tuple(o[1] for o in offs))
    for i, d in enumerate(T.data):
        idx = tuple(slice(offs[k][0][int(T.Q[k][i][0])], offs[k][0][int(T.Q[k][i][0])] + d.shape[k])
                    for k in range(T.rank))
        full[idx] += d
    return full


def _u1_layout(T, k, extra=None):
    sec = {}
    for i, d in enumerate(T.data):
        sec[int(T.Q[k][i][0])] = d.shape[k]
    if extra:
        sec.update(extra)
    off, o = {}, 0
    for q in sorted(sec):
        off[q] = o
        o += sec[q]
    return off, o, sec


def test_u1_against_dense_tensordot():
    A = gen.u1_mps(seed=41)
    C = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    la = [_u1_layout(A, k) for k in range(3)]
    FA = _dense_u1(A, la)
    ref = np.tensordot(FA, FA, axes=([0, 1], [0, 1]))
    FC = _dense_u1(C, [la[2], la[2]])
    assert np.abs(FC - ref).max() < 1e-12 * np.abs(ref).max()


# ---------------------------------------------------------------- label engine
def _brute_pairs(A, ia, B, ib):
    """all (ia,ib) with equal contracted labels; order: QC record, then the
    contracted record, then ia, then ib (the reference's stable sort of the
    sorted-merge cross product)."""
    QA, QB = A.QIDX, B.QIDX
    q = A.qdim

    def sub(Qm, i, legs):
        return tuple(int(x) for l in legs for x in Qm[i][l * q:(l + 1) * q])
    ka = [k for k in range(A.rank) if k not in ia]
    kb = [k for k in range(B.rank) if k not in ib]
    out = []
    for a in range(A.nblk):
        for b in range(B.nblk):
            if sub(QA, a, ia) == sub(QB, b, ib):
                out.append((sub(QA, a, ka) + sub(QB, b, kb), sub(QA, a, ia), a, b))
    out.sort()
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_label_engine_brute_force(seed):
    rng = np.random.RandomState(seed)
    n = 60

    def rand_t(r):
        recs = set()
        while len(recs) < n:
            recs.add(tuple(int(x) for x in rng.randint(-2, 3, size=r)))
        recs = sorted(recs, key=lambda _: rng.rand())
        Q = [np.array([[rec[k]] for rec in recs]) for k in range(r)]
        return QSpace(Q, [np.ones((1,) * r) for _ in recs], "A")
    A, B = rand_t(3), rand_t(3)
    ia, ib = [2, 0], [0, 1]
    Ia, Ib, af, seg, QC = O.pairs(A, ia, B, ib)
    ref = _brute_pairs(A, ia, B, ib)
    assert [(int(a), int(b)) for a, b in zip(Ia, Ib)] == [(r[2], r[3]) for r in ref]
    uniq = sorted(set(r[0] for r in ref))
    assert [tuple(x) for x in QC] == uniq
    assert len(seg) == len(uniq) + 1


# ---------------------------------------------------------------- errors
def test_error_paths():
    A = gen.u1_mps(seed=5)
    with pytest.raises(O.OracleError) as e:
        O.contract(A, [0], A, [2])  # itag L vs R*
    assert e.value.code == 2
    with pytest.raises(O.OracleError) as e:
        O.contract(A, [0, 0], A, [0, 1])
    assert e.value.code == 1
    with pytest.raises(O.OracleError) as e:
        O.contract(A, [5], A, [0])
    assert e.value.code == 1
    B = QSpace([np.array([[1, 2]]), np.array([[1, 2]])], [np.ones((1, 1))], "A,A", ["L", "L*"])
    with pytest.raises(O.OracleError) as e:
        O.contract(A, [0], B, [1])
    assert e.value.code == 3
    D = QSpace([np.array([[1], [1]]), np.array([[0], [0]])], [np.ones((1, 1)), np.ones((1, 1))], "A", ["x", "y"])
    with pytest.raises(O.OracleError) as e:
        O.contract(D, [0], D, [0], conjA=1)
    assert e.value.code == 4
    E = QSpace([np.array([[1]]), np.array([[1]])], [np.ones((1, 1))], "SU3", ["x", "y*"])
    with pytest.raises(O.OracleError) as e:
        O.contract(E, [1], E, [0])
    assert e.value.code == 5


def test_skip_zero_and_perm():
    A = gen.u1_mps(seed=6)
    A.data[0][:] = 0.0  # one zero block: its output blocks become exact zeros
    C = O.contract(A, [1], A, [1], conjA=1, perm=[1, 0])
    C0 = O.contract(A, [1], A, [1], conjA=1, perm=[1, 0], skip_zero=0)
    assert C.nblk < C0.nblk
    for d in C.data:
        assert np.any(d != 0)
    assert C.itags == ["R", "L*", "L", "R*"]


# ---------------------------------------------------------------- complex data
def _complexify(T, seed, imag_only_some=False):
    """T with an imaginary part drawn per block (every other block left real
    when imag_only_some)"""
    rng = np.random.RandomState(seed)
    data = []
    for i, d in enumerate(T.data):
        if imag_only_some and i % 2:
            data.append(np.asarray(d, dtype=np.complex128))
        else:
            data.append(np.asfortranarray(d + 1j * rng.standard_normal(d.shape)))
    return QSpace(T.Q, data, T.qtype, T.itags, T.info.get("cgw"))


def _conj_data(T):
    return QSpace(T.Q, [np.conj(d) for d in T.data], T.qtype, T.itags, T.info.get("cgw"))


def test_complex_u1_overlap_full_space():
    """complex U(1) MPS overlap conj(A).A and A.conj(B) against the dense
    contraction of the full tensors (the conj flag conjugates the data)"""
    A = _complexify(gen.u1_mps(seed=41, Lq=(-2, 0, 2), Rq=(-3, -1, 1, 3), dmin=2, dmax=4), 1)
    rng = np.random.RandomState(6)
    B = _complexify(QSpace(A.Q, [rng.standard_normal(d.shape) for d in A.data], A.qtype, A.itags), 2, True)
    C = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert C.is_complex()
    assert check_contraction(_conj_data(A), [0, 1], A, [0, 1], C) < 1e-13
    C2 = O.contract(A, [2], B, [2], conjB=1)
    assert check_contraction(A, [2], _conj_data(B), [2], C2) < 1e-13


def test_complex_su2_full_space():
    """complex SU(2) L.A and X.conj(A) against the dense contraction"""
    A, L = gen.su2_mps(seed=43, tS=tuple(range(5)), dmin=2, dmax=4)
    A = _complexify(A, 3)
    L = _complexify(L, 4, True)
    X = O.contract(L, [1], A, [0])
    assert check_contraction(L, [1], A, [0], X) < 1e-13
    C = O.contract(X, [0, 1], A, [0, 1], conjB=1)
    assert check_contraction(X, [0, 1], _conj_data(A), [0, 1], C) < 1e-13


def test_complex_missing_conj_rule():
    """contracted legs of equal direction and no conj flag: the reference
    applies conj(A), to the data too"""
    A = _complexify(gen.u1_mps(seed=44, Lq=(-2, 0, 2), Rq=(-3, -1, 1, 3), dmin=2, dmax=3), 5)
    C = O.contract(A, [0, 1], A, [0, 1])
    R = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert C.itags == R.itags
    for a, b in zip(C.data, R.data):
        assert np.array_equal(a, b)


def test_complex_real_imag_zero_matches_real():
    """zero imaginary parts: the complex path reproduces the real result
    bit for bit; a purely imaginary A gives i times the real contraction"""
    A, L = gen.su2_mps(seed=45, tS=tuple(range(5)), dmin=2, dmax=4)
    Ac = QSpace(A.Q, [np.asarray(d, dtype=np.complex128) for d in A.data], A.qtype, A.itags, A.info.get("cgw"))
    R = O.contract(L, [1], A, [0])
    C = O.contract(L, [1], Ac, [0])
    assert C.nblk == R.nblk and np.array_equal(C.QIDX, R.QIDX)
    for a, b in zip(C.data, R.data):
        assert np.array_equal(a.real, b) and not np.any(a.imag)
    Ai = QSpace(A.Q, [1j * np.asarray(d) for d in A.data], A.qtype, A.itags, A.info.get("cgw"))
    C2 = O.contract(Ai, [0, 1], A, [0, 1], conjA=1)
    R2 = O.contract(A, [0, 1], A, [0, 1], conjA=1)
    assert C2.nblk == R2.nblk > 0
    for a, b in zip(C2.data, R2.data):
        assert np.array_equal(a.imag, -b) and not np.any(a.real)


def test_empty_result_keeps_weight_columns():
    """a contraction whose result has no blocks (two U(1) labels per leg):
    the tensor is built, with a (0, nsym) weight array"""
    rng = np.random.RandomState(0)
    A = QSpace([np.array([[1, 0]]), np.array([[1, 0]])], [rng.standard_normal((2, 3))], "A,A", ["a", "b*"])
    B = QSpace([np.array([[2, 0]]), np.array([[2, 0]])], [rng.standard_normal((3, 2))], "A,A", ["b", "c*"])
    C = O.contract(A, [1], B, [0])
    assert C.nblk == 0 and C.rank == 2
    E = QSpace(C.Q, [], "A,A", C.itags, np.zeros((0, 2)))
    assert E.info["cgw"].shape == (0, 2)
