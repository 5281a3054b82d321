"""Full (M-resolved) space expansion of a QSpace, for convention-independent
checks: Full(T) = sum over blocks of data (x) (cgw * CGT), each leg laid out
sector by sector (ascending label record), multiplet index major, m minor.
A contraction is correct iff Full(C) = tensordot(Full(A), Full(B)).

The CGT of a block comes from the oracle (block_cgt; for outer
multiplicity the sum over the om basis tensors of om.om_basis with the
block's trailing data dimension as coefficients); any consistent CGT
convention passes, so this pins the X-symbols and the cgw/normalisation
bookkeeping, not the sign convention (pinned separately).  At most one SU(2)
symmetry per qtype (plus abelian ones).
"""
import numpy as np

import om as OM
import oracle as O


def _syms(T):
    if not T.qtype:
        return ["A"] * T.qdim
    return [s.strip() for s in T.qtype.split(",")]


def _dirs(T):
    return [(-1 if t.endswith("*") else 1) for t in T.itags] if T.itags else [1] * T.rank


def _sectors(T, k, j):
    sec = {}
    for i, d in enumerate(T.data):
        rec = tuple(int(x) for x in T.Q[k][i])
        dim = rec[j] + 1 if j is not None else 1
        sec[rec] = (d.shape[k], dim)
    return sec


def _layout(sec):
    off, o = {}, 0
    for rec in sorted(sec):
        off[rec] = o
        o += sec[rec][0] * sec[rec][1]
    return off, o


def _kron_legs(a, b):
    r = a.ndim
    out = np.multiply.outer(a, b).transpose([x for k in range(r) for x in (k, r + k)])
    return out.reshape(tuple(a.shape[k] * b.shape[k] for k in range(r)))


def expand(T, layouts):
    su2 = [j for j, s in enumerate(_syms(T)) if s == "SU2"]
    assert len(su2) <= 1
    j = su2[0] if su2 else None
    r = T.rank
    full = np.zeros(tuple(L[1] for L in layouts), dtype=np.result_type(np.float64, *[d.dtype for d in T.data]))
    dirs = _dirs(T)
    cgw = T.info.get("cgw")
    for i, d in enumerate(T.data):
        recs = [tuple(int(x) for x in T.Q[k][i]) for k in range(r)]
        if j is not None:
            tj = [rec[j] for rec in recs]
            w = cgw[i, j] if cgw is not None else O.norm_ext(tj, dirs)
            if r > 3 or np.ndim(d) > r:  # outer multiplicity: sum over the basis
                blk = OM.expand_block(d, r, tj, dirs, w)
                idx = tuple(slice(layouts[k][0][recs[k]], layouts[k][0][recs[k]] + blk.shape[k]) for k in range(r))
                full[idx] += blk
                continue
            cg = O.block_cgt(tj, dirs) * w
        else:
            cg = np.ones((1,) * r)
        blk = _kron_legs(np.asarray(d).reshape(d.shape[:r]), cg.reshape(cg.shape[:r]) if r else cg)
        idx = tuple(slice(layouts[k][0][recs[k]], layouts[k][0][recs[k]] + blk.shape[k]) for k in range(r))
        full[idx] += blk
    return full


def check_contraction(A, ia, B, ib, C):
    """max abs / scale of Full(C) - tensordot(Full(A), Full(B)); layouts of
    contracted legs from the union of both operands' sectors."""
    su2A = [j for j, s in enumerate(_syms(A)) if s == "SU2"]
    j = su2A[0] if su2A else None
    LA = [_sectors(A, k, j) for k in range(A.rank)]
    LB = [_sectors(B, k, j) for k in range(B.rank)]
    for a, b in zip(ia, ib):
        u = dict(LA[a])
        u.update(LB[b])
        LA[a] = u
        LB[b] = u
    layA = [_layout(s) for s in LA]
    layB = [_layout(s) for s in LB]
    FA = expand(A, layA)
    FB = expand(B, layB)
    ref = np.tensordot(FA, FB, axes=(list(ia), list(ib)))
    ka = [k for k in range(A.rank) if k not in ia]
    kb = [k for k in range(B.rank) if k not in ib]
    layC = [layA[k] for k in ka] + [layB[k] for k in kb]
    FC = expand(C, layC)
    scale = max(1.0, float(np.abs(ref).max()))
    return float(np.abs(FC - ref).max()) / scale
