"""Outer multiplicity in the CPU oracle (numpy restatement).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  Only tests/ import
this module.

SU(2) blocks of rank > 3 have a CGT space of dimension om >= 1; their data
carry a trailing data dimension of size om (the extra data dimension of the
reference's QS_FULL_OM blocks: ExpandOM / FuseOM / tensorRefs in
Source/QSpace_aux.cc:242-345, matrix-valued cgw in
Source/clebsch_aux.cc:4464-4471).  The basis tensors follow the library's
fusion-chain convention (stated in qspace_amd/csrc/su2.cc and restated in
om_basis below), built here from the oracle's own lowering-operator CG
tables by dense tensor contractions; the library builds them from exact
Racah coefficients by sparse recursion.

What pins each function (tests/test_oracle_om.py):
  * om_basis     -- orthonormality; invariance under the total SU(2)
                    generators (dual representation on outgoing legs);
                    completeness: om = dimension of the invariant subspace
                    of the brute-force generator matrices; rank <= 3 equals
                    the oracle's block CGT (pinned separately)
  * contract     -- full M-resolved space: Full(C) = tensordot(Full(A),
                    Full(B)) with Full built from the same basis (any
                    complete basis passes, so this pins the coefficients
                    and the data bookkeeping, not the basis convention)
  * the basis convention itself is this library's choice (the reference
    takes its OM bases from RC_STORE fusion trees, which are not here):
    parity of OM blocks with the reference's stored data is unpinned; it is
    pinned only in the full M-resolved space, which is convention-independent.
"""
from __future__ import annotations

import ctypes as C
import functools
import itertools

import numpy as np

import oracle as _o
from qspace_amd.qspace import QSpace


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


@functools.lru_cache(maxsize=None)
def cg_table(t1: int, t2: int, t3: int) -> np.ndarray:
    """<j1 m1; j2 m2 | j3 m3> of the oracle's lowering-operator construction,
    array (2j1+1, 2j2+1, 2j3+1) with index i = j - m."""
    out = np.zeros((t1 + 1) * (t2 + 1) * (t3 + 1))
    _o.lib().oracle_cg_table(t1, t2, t3, out.ctypes.data_as(C.POINTER(C.c_double)))
    return out.reshape((t1 + 1, t2 + 1, t3 + 1), order="F")


def canon_order(tj, dirs):
    r = len(tj)
    t = _o._ivec(tj)
    d = _o._ivec(dirs)
    o = np.zeros(max(1, r), np.int32)
    _o.lib().oracle_canon_order(r, _ip(t), _ip(d), _ip(o))
    return [int(x) for x in o[:r]]


def _paths(tjs):
    """intermediate couplings (J_1 = j_1, J_2, ..., J_n) of a chain, each
    level ascending, lexicographic"""
    out = [[tjs[0]]]
    for t in tjs[1:]:
        out = [p + [J] for p in out for J in range(abs(p[-1] - t), p[-1] + t + 1, 2)]
    return out


def _chain(tjs, path):
    """dense chain coupling: array (2j_1+1, ..., 2j_n+1, 2J_n+1)"""
    T = np.eye(tjs[0] + 1)
    for i in range(1, len(tjs)):
        T = np.tensordot(T, cg_table(path[i - 1], tjs[i], path[i]), axes=([T.ndim - 1], [0]))
    return T


@functools.lru_cache(maxsize=None)
def _om_basis(tj: tuple, dirs: tuple) -> np.ndarray:
    r = len(tj)
    dims = tuple(t + 1 for t in tj)
    if r == 0:
        return np.ones((1,))
    if r == 1:
        return np.ones((1, 1)) if tj[0] == 0 else np.zeros((0,) + dims)
    ord_ = canon_order(tj, dirs)
    nin = sum(1 for d in dirs if d > 0)
    flip = nin < (r + 1) // 2
    fl = [((0 if dirs[b] > 0 else 1) ^ (1 if flip else 0)) for b in ord_]
    n0 = 1
    while n0 < r and fl[n0] == fl[0]:
        n0 += 1
    ct = [tj[b] for b in ord_]
    basis = []
    if n0 < r:
        for p0 in _paths(ct[:n0]):
            for p1 in _paths(ct[n0:]):
                if p1[-1] != p0[-1]:
                    continue
                basis.append(np.tensordot(_chain(ct[:n0], p0), _chain(ct[n0:], p1), axes=([n0], [r - n0])))
    else:
        tl = ct[-1]
        sign = np.array([(-1.0) ** i for i in range(tl + 1)])
        for p0 in _paths(ct[:-1]):
            if p0[-1] != tl:
                continue
            T = _chain(ct[:-1], p0)[..., ::-1]  # last index: M = -m_last
            basis.append(T * sign)
    out = []
    for Cc in basis:
        nrm = np.sqrt(np.sum(Cc * Cc))
        if nrm == 0:
            continue
        flat = Cc.ravel(order="F")
        first = flat[np.abs(flat) > 1e-14 * nrm][0]
        Cc = Cc / (nrm if first > 0 else -nrm)
        pos = [ord_.index(b) for b in range(r)]
        out.append(np.transpose(Cc, pos))
    return np.stack(out) if out else np.zeros((0,) + dims)


def om_basis(tj, dirs) -> np.ndarray:
    """the om unit basis tensors of a block's SU(2) CGT space, array (om,
    2j_1+1, ..., 2j_r+1) in the block's leg order (library convention):
    canonical leg order; group 0 = leading canonical legs of the first flag,
    group 1 = the rest; each group coupled left to right by CG coefficients
    and the two totals identified, or (one group) legs 1..r-1 coupled to
    (j_r, -m_r) with (-1)^(j_r - m_r); paths lexicographic; unit norm with the
    first entry above 1e-14 of the norm, column-major over the canonical
    legs, positive."""
    return _om_basis(tuple(int(t) for t in tj), tuple(int(d) for d in dirs))


def xom(tja, dira, tjb, dirb, ia, ib, perm=None) -> np.ndarray:
    """x[a, b, c] = <C_c | P(contract(A_a, B_b))>"""
    A = om_basis(tja, dira)
    B = om_basis(tjb, dirb)
    ka = [k for k in range(len(tja)) if k not in ia]
    kb = [k for k in range(len(tjb)) if k not in ib]
    tjc = [tja[k] for k in ka] + [tjb[k] for k in kb]
    dc = [dira[k] for k in ka] + [dirb[k] for k in kb]
    perm = list(range(len(tjc))) if perm is None else list(perm)
    Cb = om_basis([tjc[p] for p in perm], [dc[p] for p in perm])
    X = np.zeros((len(A), len(B), len(Cb)))
    for a in range(len(A)):
        for b in range(len(B)):
            T = np.tensordot(A[a], B[b], axes=(list(ia), list(ib)))
            if T.ndim:
                T = T.transpose(perm)
            X[a, b] = [float(np.sum(Cb[c] * T)) for c in range(len(Cb))]
    return X


def _syms(T):
    return _o._syms(T)


def _dirs(tags, r, conj):
    d = [(-1 if t.endswith("*") else 1) for t in tags] if tags else [1] * r
    return [-x for x in d] if conj else d


def _comp(d, r, a):
    """component a of a block (the whole block when it has no OM dim)"""
    return d[..., a] if d.ndim > r else d


def contract(A: QSpace, ia, B: QSpace, ib, perm=None, conjA=0, conjB=0, skip_zero=1) -> QSpace:
    """A.B over legs (ia, ib) for one SU(2) symmetry with outer multiplicity
    (plus abelian symmetries): blocks paired by equal contracted labels,
    output blocks by their kept labels (sorted records), data
      C[..., c] += cgw_A cgw_B / cgw_C * sum_ab X[a, b, c] A[..., a] . B[..., b]
    (contractDATA_group with tensorRefs, Source/QSpace_aux.cc:27-345)."""
    syms = _syms(A)
    su2 = [j for j, s in enumerate(syms) if s == "SU2"]
    if len(su2) != 1:
        raise _o.OracleError(5, "outer multiplicity needs exactly one SU(2) symmetry")
    js = su2[0]
    ia, ib = list(ia), list(ib)
    ra, rb = A.rank, B.rank
    dA = _dirs(A.itags, ra, conjA)
    dB = _dirs(B.itags, rb, conjB)
    for a, b in zip(ia, ib):
        if dA[a] == dB[b]:
            raise _o.OracleError(2, "contracted legs must have opposite directions")
    ka = [k for k in range(ra) if k not in ia]
    kb = [k for k in range(rb) if k not in ib]
    rc = len(ka) + len(kb)
    perm = list(range(rc)) if perm is None else list(perm) + [k for k in range(rc) if k not in perm]
    dC0 = [dA[k] for k in ka] + [dB[k] for k in kb]
    dC = [dC0[p] for p in perm]
    tagsA = [t for t in A.itags] if A.itags else [""] * ra
    tagsB = [t for t in B.itags] if B.itags else [""] * rb

    def cj(t, c):
        if not c or not t:
            return t
        return t[:-1] if t.endswith("*") else t + "*"

    tags0 = [cj(tagsA[k], conjA) for k in ka] + [cj(tagsB[k], conjB) for k in kb]
    tagsC = [tags0[p] for p in perm] if (A.itags or B.itags) else []

    def cgw(T, i, tj, dirs_stored):
        w = T.info.get("cgw")
        return float(w[i, js]) if w is not None else _o.norm_ext(tj, dirs_stored)

    dAs = _dirs(A.itags, ra, 0)
    dBs = _dirs(B.itags, rb, 0)
    recA = [tuple(tuple(int(x) for x in A.Q[k][i]) for k in range(ra)) for i in range(A.nblk)]
    recB = [tuple(tuple(int(x) for x in B.Q[k][i]) for k in range(rb)) for i in range(B.nblk)]
    out = {}
    for i, qa in enumerate(recA):
        for j, qb in enumerate(recB):
            if any(qa[a] != qb[b] for a, b in zip(ia, ib)):
                continue
            tja = [q[js] for q in qa]
            tjb = [q[js] for q in qb]
            X = xom(tja, dA, tjb, dB, ia, ib, perm)
            if not np.any(np.abs(X) > 1e-13):
                continue
            rec0 = [qa[k] for k in ka] + [qb[k] for k in kb]
            recC = tuple(rec0[p] for p in perm)
            tjc = [q[js] for q in recC]
            fac = cgw(A, i, tja, dAs) * cgw(B, j, tjb, dBs) / _o.norm_ext(tjc, dC)
            Ad = np.conj(A.data[i]) if conjA else A.data[i]
            Bd = np.conj(B.data[j]) if conjB else B.data[j]
            oc = X.shape[2]
            for a in range(X.shape[0]):
                for b in range(X.shape[1]):
                    if not np.any(X[a, b]):
                        continue
                    T = np.tensordot(_comp(Ad, ra, a), _comp(Bd, rb, b), axes=(ia, ib))
                    if T.ndim:
                        T = T.transpose(perm)
                    acc = out.get(recC)
                    if acc is None:
                        acc = np.zeros(T.shape + (oc,), dtype=np.result_type(T.dtype, np.float64))
                        out[recC] = acc
                    acc += fac * T[..., None] * X[a, b]
    keys = sorted(out)
    data, Q = [], [[] for _ in range(rc)]
    for key in keys:
        d = out[key]
        if skip_zero and not np.any(d):
            continue
        if d.shape[-1] == 1:
            d = d[..., 0]
        for k in range(rc):
            Q[k].append(list(key[k]))
        data.append(np.asfortranarray(d))
    qd = A.qdim
    Ql = [np.array(q, dtype=np.int32).reshape(-1, qd) for q in Q]
    nsym = len(syms)
    w = np.ones((len(data), nsym))
    i = 0
    for key in keys:
        if skip_zero and not np.any(out[key]):
            continue
        w[i, js] = _o.norm_ext([q[js] for q in key], dC)
        i += 1
    return QSpace(Ql, data, A.qtype, tagsC, w)


def expand_block(d, r, tj, dirs, w):
    """data (x) (w * basis) of one block: sum over components"""
    Bs = om_basis(tj, dirs)
    full = None
    for a in range(len(Bs)):
        x = _comp(np.asarray(d), r, a)
        x = x.reshape(x.shape[:r])
        blk = _kron(x, Bs[a] * w)
        full = blk if full is None else full + blk
    return full


def _kron(a, b):
    r = a.ndim
    if r == 0:
        return a * b
    out = np.multiply.outer(a, b).transpose([x for k in range(r) for x in (k, r + k)])
    return out.reshape(tuple(a.shape[k] * b.shape[k] for k in range(r)))


def om_count(tj) -> int:
    """number of SU(2) invariants in the product of the legs' irreps
    (directions do not change it): multiplicity of J = 0, by iterated
    coupling counts"""
    mult = {tj[0]: 1}
    for t in tj[1:]:
        nxt = {}
        for J, m in mult.items():
            for K in range(abs(J - t), J + t + 1, 2):
                nxt[K] = nxt.get(K, 0) + m
        mult = nxt
    return mult.get(0, 0)


def all_labels(r, maxtj):
    return itertools.product(range(maxtj + 1), repeat=r)
