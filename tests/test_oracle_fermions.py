"""Fermionic sign factors of the contraction (host only, oracle).

The reference multiplies each pair's product by afac = fA[ia] * fB[ib]
(contractDATA_group, Source/QSpace_aux.cc:112-114) -- one sign per operand
block.  Two things are pinned here:

* the indexing: a contraction with sign vectors equals the contraction of
  the operands whose blocks were multiplied by those signs beforehand
  (bilinearity), for U(1) and SU(2) operands and every conj pattern;
* the rule that turns a Hubbard left environment into those signs
  (qspace_amd.gen.fermion_parity_signs): a two-site H*psi whose left
  environment holds the bare fermion operators of the left part, with the
  signs applied, equals the Jordan-Wigner Hamiltonian of a 2-4 site spinful
  Hubbard chain in its full Fock space (dense matrices built here from the
  site operators; the parity string of the left part enters only through
  the signs).  Without signs, or with a sign of the operator channel alone,
  the brute force fails, so the check discriminates.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from helpers import rel_err
from qspace_amd import QSpace, gen

# ---------------------------------------------------------------- indexing


def _scaled(T, f):
    return QSpace(T.Q, [np.asfortranarray(x * s) for x, s in zip(T.data, f)], T.qtype, T.itags, T.info["cgw"])


@pytest.mark.parametrize("kind", ["u1", "su2"])
@pytest.mark.parametrize("conj", [(0, 0), (1, 0), (0, 1)])
def test_sign_vectors_are_per_block_factors(kind, conj):
    rng = np.random.RandomState(5)
    if kind == "u1":
        A = gen.u1_mps(seed=3, Lq=(-4, -2, 0, 2, 4), Rq=(-5, -3, -1, 1, 3, 5), dmin=2, dmax=5)
        B = A
        ia, ib = [0, 1], [0, 1]
        if conj == (0, 0):
            conj = (1, 0)  # A.A needs one conj for the directions to match
    else:
        A, L = gen.su2_mps(seed=4, tS=tuple(range(6)), dmin=2, dmax=5)
        B = A
        ia, ib = [0, 1], [0, 1]
        if conj == (0, 0):
            A, B, ia, ib = L, A, [1], [0]
    fA = rng.choice([-1.0, 1.0], A.nblk)
    fB = rng.choice([-1.0, 1.0], B.nblk)
    got = O.contract(A, ia, B, ib, conjA=conj[0], conjB=conj[1], fermA=fA, fermB=fB, skip_zero=0)
    ref = O.contract(_scaled(A, fA), ia, _scaled(B, fB), ib, conjA=conj[0], conjB=conj[1], skip_zero=0)
    assert np.array_equal(got.QIDX, ref.QIDX)
    for a, b in zip(got.data, ref.data):
        assert np.array_equal(a, b)  # exact: a sign flip commutes with every rounding
    # a plausible indexing mistake (the signs of B used for A) changes the result
    if A is not B and A.nblk == B.nblk and not np.array_equal(fA, fB):
        bad = O.contract(A, ia, B, ib, conjA=conj[0], conjB=conj[1], fermA=fB, fermB=fA, skip_zero=0)
        assert rel_err(bad, ref) > 1e-3


# ---------------------------------------------------------------- Fock space

# one site: |0>, |up>, |dn>, |up dn> = c+_up c+_dn |0>; modes ordered up, dn
_CDAG = {
    "up": np.array([[0, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 1, 0]], float),
    "dn": np.array([[0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0], [0, -1, 0, 0]], float),
}
_Z = np.diag([1.0, -1.0, -1.0, 1.0])
_NSITE = np.array([0, 1, 1, 2])
_SZ2 = np.array([0, 1, -1, 0])  # 2 Sz


def _kron(ops):
    out = np.ones((1, 1))
    for o in ops:
        out = np.kron(out, o)
    return out


def _cdag(n, j, s):
    """c+_{j s} on n sites with the Jordan-Wigner string of sites < j"""
    return _kron([_Z] * j + [_CDAG[s]] + [np.eye(4)] * (n - j - 1))


def _hubbard(n, t, U):
    H = np.zeros((4 ** n, 4 ** n))
    for j in range(n):
        nu = _cdag(n, j, "up") @ _cdag(n, j, "up").T
        nd = _cdag(n, j, "dn") @ _cdag(n, j, "dn").T
        H += U * nu @ nd
        if j + 1 < n:
            for s in ("up", "dn"):
                h = _cdag(n, j, s) @ _cdag(n, j + 1, s).T
                H += -t * (h + h.T)
    return H


def _numbers(n):
    """particle number and 2 Sz of every basis state of n sites"""
    N = np.zeros(4 ** n, int)
    S = np.zeros(4 ** n, int)
    for idx, st in enumerate(itertools.product(range(4), repeat=n)):
        N[idx] = sum(_NSITE[x] for x in st)
        S[idx] = sum(_SZ2[x] for x in st)
    return N, S


def test_site_operators_are_fermions():
    """the dense operators obey the canonical anticommutation relations on
    3 sites (the Jordan-Wigner construction above is right), and the
    two-site Hubbard model has the textbook singlet ground state energy
    (U - sqrt(U^2 + 16 t^2)) / 2 at half filling"""
    n = 3
    ops = [_cdag(n, j, s) for j in range(n) for s in ("up", "dn")]
    I = np.eye(4 ** n)
    for a, b in itertools.product(range(len(ops)), repeat=2):
        x, y = ops[a], ops[b]
        assert np.abs(x @ y + y @ x).max() < 1e-14
        assert np.abs(x @ y.T + y.T @ x - (I if a == b else 0)).max() < 1e-14
    H = _hubbard(2, 1.0, 4.0)
    N, _ = _numbers(2)
    sec = np.where(N == 2)[0]
    e = np.linalg.eigvalsh(H[np.ix_(sec, sec)])
    assert abs(e[0] - (4.0 - np.sqrt(16.0 + 16.0)) / 2) < 1e-12


def test_parity_of_su2_labels():
    """(-1)^N = (-1)^(2S) on every common eigenstate of N and S^2 of 1-3
    sites: the parity the sign rule reads from the 2S label of an SU(2)xU(1)
    Hubbard state"""
    for n in (1, 2, 3):
        N, _ = _numbers(n)
        sp = [_cdag(n, j, "up") @ _cdag(n, j, "dn").T for j in range(n)]  # S+_j
        Sp = sum(sp)
        Sz = 0.5 * sum(_cdag(n, j, "up") @ _cdag(n, j, "up").T - _cdag(n, j, "dn") @ _cdag(n, j, "dn").T
                       for j in range(n))
        S2 = Sp.T @ Sp + Sz @ Sz + Sz
        for q in range(2 * n + 1):
            sec = np.where(N == q)[0]
            ev = np.linalg.eigvalsh(S2[np.ix_(sec, sec)])
            twoS = np.rint(np.sqrt(1.0 + 4.0 * ev) - 1.0).astype(int)  # S(S+1) -> 2S
            assert np.allclose(0.25 * twoS * (twoS + 2), ev, atol=1e-9)
            assert np.all(twoS % 2 == q % 2)


def _sectors(N, S):
    """states grouped by (N, 2Sz): {label: indices}"""
    out = {}
    for i, key in enumerate(zip(N, S)):
        out.setdefault((int(key[0]), int(key[1])), []).append(i)
    return {k: np.array(v) for k, v in sorted(out.items())}


def _two_site_operands(n, l, t, U, Ntot, Sztot, hconj):
    """HL, psi structure, HR of the cut after l of n sites: HL(a, a', w)
    holds the BARE left operators (no parity string), HR(b, b', w) the right
    ones; labels (N, 2Sz) for the left part, (N_R - Ntot, 2Sz_R - Sztot)
    for the right part (psi's blocks then have a + b = 0, the convention of
    gen.hubbard_hpsi), w = a' - a.  With hconj the hopping appears once
    (c+_L . c_R) and the local terms at half weight, the Hermitian-conjugate
    term supplying the rest."""
    r = n - l
    HLloc, HRloc = _hubbard(l, t, U), _hubbard(r, t, U)
    cl = {s: _cdag(l, l - 1, s) for s in ("up", "dn")}  # c+ of the last left site (its own JW string)
    cr = {s: _cdag(r, 0, s) for s in ("up", "dn")}      # c+ of the first right site
    half = 0.5 if hconj else 1.0
    # channels: (left operator, right operator, w label of the left operator)
    ch = [(half * HLloc, np.eye(4 ** r), (0, 0)), (np.eye(4 ** l), half * HRloc, (0, 0))]
    for s, dz in (("up", 1), ("dn", -1)):
        # c+_L c_R = (c+_L P_L) (x) c_R ; the sign rule supplies P_L
        ch.append((cl[s], -t * cr[s].T, (-1, -dz)))
        if not hconj:
            # c+_R c_L = -(c_L P_L) (x) c+_R
            ch.append((cl[s].T, t * cr[s], (1, dz)))
    NL, SL = _numbers(l)
    NR, SR = _numbers(r)
    secL = _sectors(NL, SL)
    secR = _sectors(NR - Ntot, SR - Sztot)
    wlabels = sorted({c[2] for c in ch})
    wch = {w: [c for c in ch if c[2] == w] for w in wlabels}

    def env(sec, pick, rule):
        Q, data = [[], [], []], []
        for a, ia in sec.items():
            for b, ib in sec.items():
                for w in wlabels:
                    if not rule(a, b, w):
                        continue
                    blk = np.stack([pick(c)[np.ix_(ia, ib)] for c in wch[w]], axis=2)
                    if not np.any(blk):
                        continue
                    Q[0].append(list(a)); Q[1].append(list(b)); Q[2].append(list(w))
                    data.append(np.asfortranarray(blk))
        return [np.array(q) for q in Q], data

    add = lambda x, y: (x[0] + y[0], x[1] + y[1])
    QL, dL = env(secL, lambda c: c[0], lambda a, ap, w: add(a, w) == ap)
    HL = QSpace(QL, dL, "A,A", ["E1l", "E1l*", "w"])
    QR, dR = env(secR, lambda c: c[1], lambda b, bp, w: b == add(bp, w))
    HR = QSpace(QR, dR, "A,A", ["E2r", "E2r*", "w*"])
    pq = [[], []]
    pdims = []
    for a, ia in secL.items():
        b = (-a[0], -a[1])
        if b in secR:
            pq[0].append(list(a)); pq[1].append(list(b))
            pdims.append((len(ia), len(secR[b])))
    return HL, HR, [np.array(q) for q in pq], pdims, secL, secR


def _hpsi(HL, psi, HR, hconj, fermA):
    T = O.contract(HL, [1], psi, [0], fermA=fermA, skip_zero=0)
    terms = [O.contract(T, [1, 2], HR, [2, 1], skip_zero=0)]
    if hconj:
        T2 = O.contract(HL, [0], psi, [0], conjA=1, fermA=fermA, skip_zero=0)
        terms.append(O.contract(T2, [1, 2], HR, [2, 0], conjB=1, skip_zero=0))
    key = {tuple(r): i for i, r in enumerate(psi.QIDX)}
    out = [np.zeros_like(d) for d in psi.data]
    for Ht in terms:
        for r, d in zip(Ht.QIDX, Ht.data):
            out[key[tuple(r)]] += d
    return out


@pytest.mark.parametrize("n,l,Ntot,Sztot", [(2, 1, 2, 0), (3, 1, 3, 1), (3, 2, 2, 0), (4, 2, 4, 0), (4, 1, 3, -1),
                                            (4, 3, 5, 1)])
@pytest.mark.parametrize("hconj", [0, 1])
def test_two_site_hpsi_matches_fock_space(n, l, Ntot, Sztot, hconj):
    t, U = 1.0, 4.0
    HL, HR, pq, pdims, secL, secR = _two_site_operands(n, l, t, U, Ntot, Sztot, hconj)
    rng = np.random.RandomState(n * 10 + l)
    psi = QSpace(pq, [np.asfortranarray(rng.standard_normal(d)) for d in pdims], "A,A", ["E1l", "E2r"])
    # the dense wave function in the full Fock space (left index major)
    dimR = 4 ** (n - l)
    v = np.zeros(4 ** n)
    for r, d in zip(psi.QIDX, psi.data):
        ia, ib = secL[(r[0], r[1])], secR[(r[2], r[3])]
        v[(ia[:, None] * dimR + ib[None, :]).ravel()] = d.ravel()
    Hv = _hubbard(n, t, U) @ v
    ref = [Hv[(secL[(r[0], r[1])][:, None] * dimR + secR[(r[2], r[3])][None, :]).ravel()].reshape(d.shape)
           for r, d in zip(psi.QIDX, psi.data)]
    nrm = np.sqrt(sum(float(np.sum(x * x)) for x in ref))

    def err(out):
        return np.sqrt(sum(float(np.sum((a - b) ** 2)) for a, b in zip(out, ref))) / nrm

    signs = gen.fermion_parity_signs(HL, parity=lambda rec: int(rec[0]) % 2)
    assert err(_hpsi(HL, psi, HR, hconj, signs)) < 1e-13
    if n - l >= 1 and l >= 1 and Ntot > 1:
        # no signs, or a sign of the operator channel alone (a gauge of the
        # cut's hopping, c_R -> -c_R): not the Jordan-Wigner Hamiltonian
        chan = np.where(HL.Q[2][:, 0] % 2 == 0, 1.0, -1.0)
        assert err(_hpsi(HL, psi, HR, hconj, None)) > 1e-3
        assert err(_hpsi(HL, psi, HR, hconj, chan)) > 1e-3


def test_hubbard_workload_signs_follow_the_ket_parity():
    """the signs of the SU(2)xU(1) Hubbard workload: -1 exactly on the blocks
    of a fermionic operator channel (odd 2S of w) whose ket leg (E1l*, the
    leg contracted with psi) has odd 2S; every channel has blocks of both
    signs (the string is not a per-channel constant)"""
    HL, _, _ = gen.hubbard_hpsi(seed=26, Dstar=30, mu=1.0, sigma=0.4, dmax=4, qmax=3)
    f = gen.fermion_parity_signs(HL)
    ket, w = HL.Q[1][:, 1] % 2, HL.Q[2][:, 1] % 2
    assert np.array_equal(f, np.where(ket * w == 1, -1.0, 1.0))
    for lab in {tuple(r) for r in HL.Q[2] if r[1] % 2}:
        sel = np.all(HL.Q[2] == lab, axis=1)
        assert set(f[sel]) == {-1.0, 1.0}
