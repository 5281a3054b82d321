"""Seeded synthetic inputs for the five contraction workloads.

The reference ships no data, so every input is generated: labels and block
dims follow the workload definitions, block entries are FP64 N(0,1).  The
random stream is numpy's legacy MT19937 ``RandomState`` (frozen across numpy
versions).  SU(2) labels are 2S; U(1) labels are integer charges.

Workloads (sizes at full scale):
  u1_mps       U(1) rank-3 A(L,s,R), L charges -10..10 step 2 (dims U[4,32]),
               s = +-1 (dim 1), R charges -11..11 step 2 (dims U[4,32]);
               C = contractQS(A,'12*',A,'12').  The odd R charges are needed
               for blocks to exist (qL+qs=qR with even qL and odd qs).
  su2_env      SU(2) L(L,L') . A(L,s,R) . conj(A): 2S in 0..16 on both bonds
               (17 labels x mean dim 36 ~ 600 multiplets), dims U[8,64],
               local spin 1/2.
  heis_hpsi    two-site H*psi of the spin-1/2 Heisenberg chain in SU(2), in the
               enlarged-bond form psi(E1l,E2r) with rank-3 environments
               HK(E,E',w); MPO bond w = {2S=0: dim 2, 2S=2: dim 1} (3 multiplets);
               enlarged-bond sectors 2S = 0,2,4,... with lognormal(mu,sigma)
               dims clipped to [1,dmax], drawn until sum(d) >= D*.
  heis_pseudo_hpsi  the same chain in the pseudo-MPO form of get_HPsi.m:62-83:
               block Hamiltonians HK_L, HK_R and spin-1 operator pairs.
  hubbard_hpsi SU(2)xU(1) ("A,SU2") Hubbard two-site H*psi, MPO bond
               {(0,0): 2, (+1,1): 2, (-1,1): 2} (6 multiplets), sectors (Q,2S)
               with Q in [-12,12], Q+2S even, ordered by |Q|+2S, lognormal dims.
"""
from __future__ import annotations

import itertools

import numpy as np

from .qspace import QSpace


def _rng(seed):
    return np.random.RandomState(seed)


def _normal_block(rng, shape):
    return np.asfortranarray(rng.standard_normal(int(np.prod(shape))).reshape(shape, order="F"))


# ---------------------------------------------------------------------------
def u1_mps(seed=1001, Lq=tuple(range(-10, 11, 2)), Rq=tuple(range(-11, 12, 2)), dmin=4, dmax=32):
    """Rank-3 U(1) MPS tensor A(L,s,R) with qL + qs = qR."""
    rng = _rng(seed)
    dL = {q: int(rng.randint(dmin, dmax + 1)) for q in Lq}
    dR = {q: int(rng.randint(dmin, dmax + 1)) for q in Rq}
    Q = [[], [], []]
    data = []
    for qL in Lq:
        for qs in (-1, 1):
            qR = qL + qs
            if qR not in dR:
                continue
            Q[0].append([qL]); Q[1].append([qs]); Q[2].append([qR])
            data.append(_normal_block(rng, (dL[qL], 1, dR[qR])))
    return QSpace([np.array(q) for q in Q], data, "A", ["L", "s", "R*"])


# ---------------------------------------------------------------------------
def su2_labels_dims(rng, labels, dmin, dmax):
    return {q: int(rng.randint(dmin, dmax + 1)) for q in labels}


def su2_mps(seed=1002, tS=tuple(range(0, 17)), dmin=8, dmax=64):
    """Rank-3 SU(2) A(L,s,R): L in, s (2S=1) in, R out; blocks |S_L-S_R|=1/2."""
    rng = _rng(seed)
    dL = su2_labels_dims(rng, tS, dmin, dmax)
    dR = su2_labels_dims(rng, tS, dmin, dmax)
    Q = [[], [], []]
    data = []
    for a in tS:
        for b in tS:
            if abs(a - b) != 1:
                continue
            Q[0].append([a]); Q[1].append([1]); Q[2].append([b])
            data.append(_normal_block(rng, (dL[a], 1, dR[b])))
    A = QSpace([np.array(q) for q in Q], data, "SU2", ["L", "s", "R*"])
    L = QSpace([np.array([[a] for a in tS]), np.array([[a] for a in tS])],
               [_normal_block(rng, (dL[a], dL[a])) for a in tS], "SU2", ["L", "L*"])
    return A, L


# ---------------------------------------------------------------------------
def lognormal_sectors(rng, labels, Dstar, mu, sigma, dmax):
    """Assign lognormal dims round(exp(mu + sigma Z)) clipped to [1,dmax] to
    labels in order until the dims sum to at least Dstar."""
    out = []
    tot = 0
    for q in labels:
        if tot >= Dstar:
            break
        d = int(np.clip(np.rint(np.exp(mu + sigma * rng.standard_normal())), 1, dmax))
        out.append((q, d))
        tot += d
    if tot < Dstar:
        raise ValueError("label set too small for D*")
    return out


def _env_blocks(rng, secA, secB, wsec, rule, scale=1.0):
    """rank-3 environment blocks (a, a', w) for the label rule."""
    Q = [[], [], []]
    data = []
    for qa, da in secA:
        for qb, db in secB:
            for qw, dw in wsec:
                if rule(qa, qb, qw):
                    Q[0].append(list(np.atleast_1d(qa)))
                    Q[1].append(list(np.atleast_1d(qb)))
                    Q[2].append(list(np.atleast_1d(qw)))
                    data.append(scale * _normal_block(rng, (da, db, dw)))
    return Q, data


def _tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def heis_hpsi(seed=1003, Dstar=2000, mu=3.0, sigma=0.8, dmax=256):
    """SU(2) Heisenberg two-site H*psi operands: (HL, psi, HR)."""
    rng = _rng(seed)
    labels = list(range(0, 4000, 2))
    E1 = lognormal_sectors(rng, labels, Dstar, mu, sigma, dmax)
    E2 = [(q, int(np.clip(np.rint(np.exp(mu + sigma * rng.standard_normal())), 1, dmax))) for q, _ in E1]
    wsec = [(0, 2), (2, 1)]
    rule = lambda a, b, w: _tri(a, b, w)
    QL, dL = _env_blocks(rng, E1, E1, wsec, rule)
    HL = QSpace([np.array(q) for q in QL], dL, "SU2", ["E1l", "E1l*", "w"])
    QR, dR = _env_blocks(rng, E2, E2, wsec, rule)
    HR = QSpace([np.array(q) for q in QR], dR, "SU2", ["E2r", "E2r*", "w*"])
    psi = QSpace([np.array([[q] for q, _ in E1]), np.array([[q] for q, _ in E2])],
                 [_normal_block(rng, (d1, d2)) for (_, d1), (_, d2) in zip(E1, E2)], "SU2", ["E1l", "E2r"])
    nrm = psi.norm()
    psi = psi.with_packed(psi.packed() / nrm)
    return HL, psi, HR


def heis_pseudo_hpsi(seed=1008, Dstar=2000, mu=3.0, sigma=0.8, dmax=256, nop=1):
    """Operands of the pseudo-MPO two-site H*psi of the SU(2) spin-1/2
    Heisenberg chain (get_HPsi.m:62-83): psi(E1l,E2r) as in heis_hpsi, the
    block Hamiltonians HK_L(E1l,E1l*) and HK_R(E2r,E2r*) (symmetric blocks),
    and nop spin-1 operator pairs OP_L(E1l,E1l*,op_j), OP_R(E2r,E2r*,op_j)
    (operator leg 2S = 2, incoming on both, matched by its itag).
    Returns (HKL, psi, HKR, [OP_L], [OP_R])."""
    rng = _rng(seed)
    labels = list(range(0, 4000, 2))
    E1 = lognormal_sectors(rng, labels, Dstar, mu, sigma, dmax)
    E2 = [(q, int(np.clip(np.rint(np.exp(mu + sigma * rng.standard_normal())), 1, dmax))) for q, _ in E1]

    def block_h(E, tags):
        data = []
        for _, d in E:
            x = rng.standard_normal((d, d))
            data.append(np.asfortranarray(0.5 * (x + x.T)))
        return QSpace([np.array([[q] for q, _ in E]), np.array([[q] for q, _ in E])], data, "SU2", tags)

    HKL = block_h(E1, ["E1l", "E1l*"])
    HKR = block_h(E2, ["E2r", "E2r*"])
    rule = lambda a, b, w: _tri(a, b, w)
    OPL, OPR = [], []
    for j in range(nop):
        tag = f"op{j + 1}"
        QL, dL = _env_blocks(rng, E1, E1, [(2, 1)], rule)
        OPL.append(QSpace([np.array(q) for q in QL], dL, "SU2", ["E1l", "E1l*", tag]))
        QR, dR = _env_blocks(rng, E2, E2, [(2, 1)], rule)
        OPR.append(QSpace([np.array(q) for q in QR], dR, "SU2", ["E2r", "E2r*", tag]))
    psi = QSpace([np.array([[q] for q, _ in E1]), np.array([[q] for q, _ in E2])],
                 [_normal_block(rng, (d1, d2)) for (_, d1), (_, d2) in zip(E1, E2)], "SU2", ["E1l", "E2r"])
    psi = psi.with_packed(psi.packed() / psi.norm())
    return HKL, psi, HKR, OPL, OPR


def hubbard_hpsi(seed=1005, Dstar=4000, mu=3.2, sigma=0.9, dmax=384, qmax=12):
    """SU(2)xU(1) Hubbard two-site H*psi operands: (HL, psi, HR) with qtype 'A,SU2'."""
    rng = _rng(seed)
    cand = [(Q, s) for Q in range(-qmax, qmax + 1) for s in range(0, 64) if (Q + s) % 2 == 0]
    cand.sort(key=lambda x: (abs(x[0]) + x[1], x[1], x[0]))
    E1 = lognormal_sectors(rng, cand, Dstar, mu, sigma, dmax)
    E2 = [((-Q, s), int(np.clip(np.rint(np.exp(mu + sigma * rng.standard_normal())), 1, dmax))) for (Q, s), _ in E1]
    wsec = [((0, 0), 2), ((1, 1), 2), ((-1, 1), 2)]
    # HL(a in, a' out, w in): Q_a + Q_w = Q_a'; HR(b in, b' out, w out): Q_b = Q_b' + Q_w
    ruleL = lambda a, b, w: a[0] + w[0] == b[0] and _tri(a[1], b[1], w[1])
    ruleR = lambda a, b, w: a[0] == b[0] + w[0] and _tri(a[1], b[1], w[1])
    QL, dL = _env_blocks(rng, E1, E1, wsec, ruleL)
    HL = QSpace([np.array(q) for q in QL], dL, "A,SU2", ["E1l", "E1l*", "w"])
    QR, dR = _env_blocks(rng, E2, E2, wsec, ruleR)
    HR = QSpace([np.array(q) for q in QR], dR, "A,SU2", ["E2r", "E2r*", "w*"])
    psi = QSpace([np.array([list(q) for q, _ in E1]), np.array([list(q) for q, _ in E2])],
                 [_normal_block(rng, (d1, d2)) for (_, d1), (_, d2) in zip(E1, E2)], "A,SU2", ["E1l", "E2r"])
    nrm = psi.norm()
    psi = psi.with_packed(psi.packed() / nrm)
    return HL, psi, HR


def fermion_parity_signs(HL: QSpace, ket_leg: int = 1, w_leg: int = 2, parity=None) -> np.ndarray:
    """Exact +-1 per block of a left environment HL(a, a', w) that holds the
    bare fermion operators of the left part: the Jordan-Wigner string P_L.

    A fermion operator of the right part is P_L (x) c_R, so the cut's
    hopping c+_L c_R = (c+_L P_L) (x) c_R and c+_R c_L = -(c_L P_L) (x) c+_R:
    the left operator of every fermionic channel (odd parity of its w label)
    is followed by P_L, i.e. multiplied by the parity (-1)^N of its ket
    state a' (ket_leg, the leg contracted with psi).  The sign is applied
    per block through the contraction's afac = fA[ia] * fB[ib]
    (Source/QSpace_aux.cc:112-114), and in the Hermitian-conjugate term the
    same block represents (c+_L P_L)^T = P_L c_L with the same factor.
    ``parity`` maps a leg's label record to 0/1; the default reads the 2S
    of an "A,SU2" Hubbard label (Q, 2S): a block of sites has (-1)^N =
    (-1)^(2S), since only singly occupied sites carry spin 1/2.  Both the
    rule and the parity map are checked against a full Fock-space
    Jordan-Wigner Hamiltonian of 2-4 sites in the tests."""
    if parity is None:
        parity = lambda rec: int(rec[1]) % 2
    z = np.array([parity(HL.Q[w_leg][i]) * parity(HL.Q[ket_leg][i]) for i in range(HL.nblk)], dtype=np.int64)
    return np.where(z % 2 == 0, 1.0, -1.0)


# ---------------------------------------------------------------------------
def u1_env_update(seed=1006, Lq=tuple(range(-6, 7, 2)), Rq=tuple(range(-7, 8, 2)), dmin=3, dmax=10,
                  ws=(-2, 0, 2), dw=1):
    """Operands of the U(1) environment update L' = A'.(HK.W.A) in the leg
    conventions of updateHK.m: A(L,R,s) (reference leg order), HK(L*,L,w*)
    and the MPO W(w,w2*,s*,s), charges conserved (in = out)."""
    rng = _rng(seed + 7)
    A0 = u1_mps(seed=seed, Lq=Lq, Rq=Rq, dmin=dmin, dmax=dmax)
    A = QSpace([A0.Q[0], A0.Q[2], A0.Q[1]], [np.asfortranarray(np.transpose(d, (0, 2, 1))) for d in A0.data], "A",
               ["L", "R*", "s"])
    dL = {int(q): A.data[i].shape[0] for i, q in enumerate(A.Q[0][:, 0])}
    hq, hd = [[], [], []], []
    for a in sorted(dL):
        for w in ws:
            b = a + w
            if b in dL:
                hq[0].append([a]); hq[1].append([b]); hq[2].append([w])
                hd.append(_normal_block(rng, (dL[a], dL[b], dw)))
    HK = QSpace([np.array(q) for q in hq], hd, "A", ["L*", "L", "w*"])
    wq, wd = [[], [], [], []], []
    for w in ws:
        for s_in in (-1, 1):
            for w2 in ws:
                s_out = w + s_in - w2
                if s_out in (-1, 1):
                    for k, v in enumerate((w, w2, s_out, s_in)):
                        wq[k].append([v])
                    wd.append(_normal_block(rng, (dw, dw, 1, 1)))
    W = QSpace([np.array(q) for q in wq], wd, "A", ["w", "w2*", "s*", "s"])
    return A, HK, W


# ---------------------------------------------------------------------------
def u1_two_site_psi(seed=1007, Lq=tuple(range(-10, 11, 2)), Rq=tuple(range(-10, 11, 2)), dmin=4, dmax=32,
                    itags=("L", "s1", "s2*", "R*")):
    """Rank-4 U(1) two-site wave function psi(L,s1,s2,R) (spin-1/2 sites,
    charges 2Sz = +-1, dim 1) with qL + qs1 = qs2 + qR when the itags are
    (in, in, out, out); the group of legs (L,s1) meets (s2,R) in the
    truncation, which the reference requires to share one direction."""
    rng = _rng(seed)
    dL = {q: int(rng.randint(dmin, dmax + 1)) for q in Lq}
    dR = {q: int(rng.randint(dmin, dmax + 1)) for q in Rq}
    sgn = [(-1 if t.endswith("*") else 1) for t in itags]
    Q = [[], [], [], []]
    data = []
    for qL in Lq:
        for q1 in (-1, 1):
            for q2 in (-1, 1):
                for qR in Rq:
                    if sgn[0] * qL + sgn[1] * q1 + sgn[2] * q2 + sgn[3] * qR != 0:
                        continue
                    for k, v in enumerate((qL, q1, q2, qR)):
                        Q[k].append([v])
                    data.append(_normal_block(rng, (dL[qL], 1, 1, dR[qR])))
    return QSpace([np.array(q) for q in Q], data, "A", list(itags))


# ---------------------------------------------------------------------------
def _su2_invariants(tj) -> int:
    """number of SU(2) invariants of a leg set (multiplicity of total spin 0,
    by iterated coupling counts; directions do not change it)"""
    mult = {int(tj[0]): 1}
    for t in tj[1:]:
        nxt = {}
        for J, m in mult.items():
            for K in range(abs(J - int(t)), J + int(t) + 1, 2):
                nxt[K] = nxt.get(K, 0) + m
        mult = nxt
    return mult.get(0, 0)


def su2_om_tensor(seed, legs, itags, scale=1.0, keep=1.0):
    """random SU(2) QSpace of any rank: every combination of the legs'
    (2S, dim) sectors that admits an invariant becomes a block; a block
    whose CGT space has dimension om > 1 carries a trailing data dimension
    of size om (outer multiplicity).  keep < 1 drops blocks at random."""
    rng = _rng(seed)
    r = len(legs)
    Q = [[] for _ in range(r)]
    data = []
    for combo in itertools.product(*legs):
        tj = [c[0] for c in combo]
        om = _su2_invariants(tj)
        if not om or (keep < 1 and rng.rand() > keep):
            continue
        shape = tuple(c[1] for c in combo) + ((om,) if om > 1 else ())
        for k in range(r):
            Q[k].append([tj[k]])
        data.append(scale * _normal_block(rng, shape))
    return QSpace([np.array(q, dtype=np.int32).reshape(-1, 1) for q in Q], data, "SU2", list(itags))


def su2_env_update(seed=1009, tS=tuple(range(0, 7)), dmin=2, dmax=6, dw=2):
    """Operands of the SU(2) environment update L' = A'.(HK.W.A) in the leg
    conventions of updateHK.m: A(L,R*,s) (s = spin 1/2), HK(L*,L,w*) and the
    rank-4 MPO W(w,w2*,s*,s) whose operator legs carry spin 0 and 1 (the
    identity and spin-operator channels of a Heisenberg MPO), random data;
    W's blocks and the rank-4 intermediates have outer multiplicity."""
    rng = _rng(seed)
    A0, _ = su2_mps(seed=seed, tS=tS, dmin=dmin, dmax=dmax)
    A = QSpace([A0.Q[0], A0.Q[2], A0.Q[1]], [np.asfortranarray(np.transpose(d, (0, 2, 1))) for d in A0.data],
               "SU2", ["L", "R*", "s"])
    dL = {int(q): A.data[i].shape[0] for i, q in enumerate(A.Q[0][:, 0])}
    secL = sorted(dL.items())
    wsec = [(0, dw), (2, dw)]
    HK = su2_om_tensor(seed + 1, [secL, secL, wsec], ["L*", "L", "w*"])
    W = su2_om_tensor(seed + 2, [wsec, wsec, [(1, 1)], [(1, 1)]], ["w", "w2*", "s*", "s"])
    return A, HK, W
