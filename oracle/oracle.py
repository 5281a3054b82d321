"""ctypes front of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker of the HIP path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg import this module.

Parity status of the oracle's functions (what pins each one):
  * label engine (sort/match/group)  -- brute-force nested-loop matcher with an
    independently stated order on tiny inputs (tests/test_oracle.py)
  * per-pair data (U(1))             -- dense embedding + numpy.tensordot
  * SU(2) CG coefficients             -- sympy.physics.wigner.clebsch_gordan (exact)
  * SU(2) CGT normalisation/sign      -- explicit restated rule checks
  * X-symbols (x3)                    -- full M-resolved space contraction
  * SkipZeroData / Permute / NormCGW  -- closed-form checks (norm, permutation)
  * permute (permuteQS)               -- full M-resolved space: Full(permute(A))
                                         = transpose(Full(A)); conj flag vs the
                                         contraction's conj flag
  * plus (plusQS)                     -- full M-resolved space: Full(C) = Full(A)
                                         + bfac Full(B); block order and zero
                                         handling by hand-built cases
  * complex data (contract)           -- four real contractions combined by
                                         bilinearity; full M-resolved space
                                         with the conjugated operand's data
                                         conjugated, zero imaginary parts =
                                         the real result bit for bit
  * outer multiplicity (om.py)        -- see om.py: basis orthonormal,
                                         invariant and complete; block-wise
                                         contraction in the full M-resolved
                                         space
  * ortho (orthoQS)                   -- dense singular values, A1 isometry,
                                         A1.A2 = PSI (kept all) or the scaled
                                         projection / Eckart-Young (truncated),
                                         markSet and itag rules by literal
                                         cases; SU(2) K = 1, r-1 in the full
                                         M-resolved space, weights other than
                                         normExt included
  * struct encoding / decoding         -- its own (abi_io.py), pinned against
                                         the reference's itag layout by known
                                         answers and against the product's
                                         encoder byte for byte
  * reference's FP64 summation order / MKL dgemm rounding: parity unpinned
    (no reference build; tolerance 1e-12 relative instead of bit equality)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from abi_io import Tensor as QsxTensor, TensorOut as QsxTensorOut, decode as from_out, encode as to_c
from qspace_amd.qspace import QSpace

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        L.oracle_contract.argtypes = [P(QsxTensor), P(C.c_int32), C.c_int32, C.c_int32, P(QsxTensor), P(C.c_int32),
                                      C.c_int32, C.c_int32, P(C.c_int32), C.c_int32, P(C.c_double), P(C.c_double),
                                      C.c_int32, P(P(QsxTensorOut)), C.c_char_p, C.c_size_t]
        L.oracle_pairs.argtypes = [P(QsxTensor), P(C.c_int32), C.c_int32, C.c_int32, P(QsxTensor), P(C.c_int32),
                                   C.c_int32, C.c_int32, P(C.c_int64), P(C.c_int64), P(C.c_int32), P(C.c_int32),
                                   P(C.c_double), P(C.c_int64), P(C.c_int32), C.c_char_p, C.c_size_t]
        L.oracle_tensor_free.argtypes = [P(QsxTensorOut)]
        L.oracle_su2_cg.argtypes = [C.c_int32] * 6
        L.oracle_su2_cg.restype = C.c_double
        L.oracle_block_cgt.argtypes = [C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_double), C.c_int64, C.c_char_p,
                                       C.c_size_t]
        L.oracle_block_cgt.restype = C.c_int64
        L.oracle_cg_table.argtypes = [C.c_int32, C.c_int32, C.c_int32, P(C.c_double)]
        L.oracle_cg_table.restype = None
        L.oracle_canon_order.argtypes = [C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32)]
        L.oracle_canon_order.restype = None
        L.oracle_norm_ext.argtypes = [C.c_int32, P(C.c_int32), P(C.c_int32)]
        L.oracle_norm_ext.restype = C.c_double
        L.oracle_x3.argtypes = [C.c_int32, P(C.c_int32), P(C.c_int32), C.c_int32, P(C.c_int32), P(C.c_int32),
                                P(C.c_int32), P(C.c_int32), C.c_int32, P(C.c_double), C.c_char_p, C.c_size_t]
        L.oracle_load_blas.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
        L.oracle_baseline_create.argtypes = [P(QsxTensor), P(C.c_int32), C.c_int32, C.c_int32, P(QsxTensor),
                                             P(C.c_int32), C.c_int32, C.c_int32, C.c_char_p, C.c_size_t]
        L.oracle_baseline_create.restype = C.c_void_p
        L.oracle_baseline_flops.argtypes = [C.c_void_p]
        L.oracle_baseline_flops.restype = C.c_double
        L.oracle_baseline_run.argtypes = [C.c_void_p, C.c_int32]
        L.oracle_baseline_output.argtypes = [C.c_void_p, P(C.c_double), C.c_int64]
        L.oracle_baseline_output.restype = C.c_int64
        L.oracle_baseline_destroy.argtypes = [C.c_void_p]
        L.oracle_blas_name.restype = C.c_char_p
        _LIB = L
    return _LIB


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def _ivec(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1))
    if a.size == 0:
        a = np.zeros(1, np.int32)
    return a


def _needs_om(A: QSpace, ia, B: QSpace, ib) -> bool:
    if "SU2" not in _syms(A):
        return False
    rc = A.rank + B.rank - 2 * len(ia)
    return A.rank > 3 or B.rank > 3 or rc > 3 or max(A.om(), default=1) > 1 or max(B.om(), default=1) > 1


def contract(A: QSpace, ia, B: QSpace, ib, perm=None, conjA=0, conjB=0, fermA=None, fermB=None,
             skip_zero=1) -> QSpace:
    """0-based ia/ib; returns the oracle's contractQS result.  SU(2)
    contractions with outer multiplicity (an operand or the result of rank
    > 3, or blocks with an OM data dimension) go to om.contract, whose
    output blocks are in sorted record order."""
    if _needs_om(A, ia, B, ib):
        if fermA is not None or fermB is not None:
            raise OracleError(5, "fermionic signs with outer multiplicity are not restated")
        import om
        return om.contract(A, ia, B, ib, perm=perm, conjA=conjA, conjB=conjB, skip_zero=skip_zero)
    if A.is_complex() or B.is_complex():
        return _contract_complex(A, ia, B, ib, perm, conjA, conjB, fermA, fermB, skip_zero)
    L = lib()
    ca, ka = to_c(A)
    cb, kb = to_c(B)
    ia_ = _ivec(ia)
    ib_ = _ivec(ib)
    p = _ivec(perm if perm is not None else [])
    fa = None if fermA is None else np.ascontiguousarray(fermA, dtype=np.float64)
    fb = None if fermB is None else np.ascontiguousarray(fermB, dtype=np.float64)
    out = C.POINTER(QsxTensorOut)()
    err = C.create_string_buffer(512)
    rc = L.oracle_contract(C.byref(ca), ia_.ctypes.data_as(C.POINTER(C.c_int32)), len(ia), conjA, C.byref(cb),
                           ib_.ctypes.data_as(C.POINTER(C.c_int32)), len(ib), conjB,
                           p.ctypes.data_as(C.POINTER(C.c_int32)), 0 if perm is None else len(perm),
                           None if fa is None else fa.ctypes.data_as(C.POINTER(C.c_double)),
                           None if fb is None else fb.ctypes.data_as(C.POINTER(C.c_double)), skip_zero,
                           C.byref(out), err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    try:
        T = from_out(out.contents)
        T.npairs = out.contents.npairs
        T.flops = out.contents.useful_flops
    finally:
        L.oracle_tensor_free(out)
    return T


def _planes(T: QSpace):
    re = QSpace(T.Q, [np.real(d).astype(np.float64) for d in T.data], T.qtype, T.itags, T.info.get("cgw"))
    if not T.is_complex():
        return re, None
    im = QSpace(T.Q, [np.imag(d).astype(np.float64) for d in T.data], T.qtype, T.itags, T.info.get("cgw"))
    return re, im


def _contract_complex(A, ia, B, ib, perm, conjA, conjB, fermA, fermB, skip_zero) -> QSpace:
    """complex data by bilinearity.  The reference conjugates the data of an
    operand contracted with its conj flag (QSpace::contract passes the flag
    to the block products, Source/QSpace.cc:4186-4187; wbarray::contract
    conjugates a complex operand, Source/wbarray.cc:3473-3483), so with
    s = -1 for a conjugated operand
      Re C = C(Ar,Br) - sA sB C(Ai,Bi),  Im C = sB C(Ar,Bi) + sA C(Ai,Br)
    from four real contractions of the same structure (all blocks kept),
    then SkipZeroData on the combined blocks (both planes zero; rank-2
    block-diagonal results keep theirs, Source/QSpace.cc:3200-3305).
    The conj flag of the data is the one after the missing-conj rule
    (contract_getIdxSet, Source/QSpace.cc:4262-4267): contracted legs of
    equal direction with neither or both operands conjugated mean conj(A)."""
    if not (conjA ^ conjB) and A.itags and B.itags and len(ia):
        if all(A.itags[a].endswith("*") == B.itags[b].endswith("*") for a, b in zip(ia, ib)):
            conjA = 1
    Ar, Ai = _planes(A)
    Br, Bi = _planes(B)

    def one(X, Y):
        return contract(X, ia, Y, ib, perm=perm, conjA=conjA, conjB=conjB, fermA=fermA, fermB=fermB,
                        skip_zero=0)

    R1 = one(Ar, Br)
    sA = -1.0 if conjA else 1.0
    sB = -1.0 if conjB else 1.0
    re = [np.asarray(d, dtype=np.float64) for d in R1.data]
    im = [np.zeros_like(d) for d in re]
    if Ai is not None and Bi is not None:
        re = [x + (-sA * sB) * y for x, y in zip(re, one(Ai, Bi).data)]
    if Bi is not None:
        im = [sB * y for y in one(Ar, Bi).data]
    if Ai is not None:
        R4 = one(Ai, Br).data
        im = [sA * y for y in R4] if Bi is None else [x + sA * y for x, y in zip(im, R4)]
    data = [x + 1j * y for x, y in zip(re, im)]
    keep = list(range(len(data)))
    if skip_zero and data:
        rc = R1.rank
        blockdiag = rc == 2 and all(np.array_equal(R1.Q[0][i], R1.Q[1][i]) for i in range(len(data)))
        if not blockdiag:
            keep = [i for i, d in enumerate(data) if np.any(d != 0)]
    cg = R1.info.get("cgw")
    T = QSpace([q[keep] for q in R1.Q], [data[i] for i in keep], R1.qtype, R1.itags,
               None if cg is None else cg[keep])
    T.npairs = R1.npairs
    T.flops = R1.flops * ((Ai is not None and Bi is not None) + (Ai is not None) + (Bi is not None) + 1)
    return T


def pairs(A: QSpace, ia, B: QSpace, ib, conjA=0, conjB=0):
    """structure half: (Ia, Ib, afac, seg, QC records)"""
    L = lib()
    ca, ka = to_c(A, with_data=False)
    cb, kb = to_c(B, with_data=False)
    ia_ = _ivec(ia)
    ib_ = _ivec(ib)
    npair = C.c_int64()
    nout = C.c_int64()
    err = C.create_string_buffer(512)
    P = C.POINTER
    rc = L.oracle_pairs(C.byref(ca), ia_.ctypes.data_as(P(C.c_int32)), len(ia), conjA, C.byref(cb),
                        ib_.ctypes.data_as(P(C.c_int32)), len(ib), conjB, C.byref(npair), C.byref(nout), None, None,
                        None, None, None, err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    n, m = npair.value, nout.value
    rc_ = A.rank + B.rank - 2 * len(ia)
    Ia = np.zeros(max(1, n), np.int32)
    Ib = np.zeros(max(1, n), np.int32)
    af = np.zeros(max(1, n))
    seg = np.zeros(m + 1, np.int64)
    QC = np.zeros(max(1, m * rc_ * A.qdim), np.int32)
    rc = L.oracle_pairs(C.byref(ca), ia_.ctypes.data_as(P(C.c_int32)), len(ia), conjA, C.byref(cb),
                        ib_.ctypes.data_as(P(C.c_int32)), len(ib), conjB, C.byref(npair), C.byref(nout),
                        Ia.ctypes.data_as(P(C.c_int32)), Ib.ctypes.data_as(P(C.c_int32)),
                        af.ctypes.data_as(P(C.c_double)), seg.ctypes.data_as(P(C.c_int64)),
                        QC.ctypes.data_as(P(C.c_int32)), err, 512)
    if rc:
        raise OracleError(rc, err.value.decode())
    return Ia[:n], Ib[:n], af[:n], seg, QC[: m * rc_ * A.qdim].reshape(m, rc_ * A.qdim)


def su2_cg(tj1, tm1, tj2, tm2, tj3, tm3) -> float:
    return lib().oracle_su2_cg(tj1, tm1, tj2, tm2, tj3, tm3)


def block_cgt(tj, dirs) -> np.ndarray:
    L = lib()
    r = len(tj)
    t = _ivec(tj)
    d = _ivec(dirs)
    err = C.create_string_buffer(256)
    n = L.oracle_block_cgt(r, t.ctypes.data_as(C.POINTER(C.c_int32)), d.ctypes.data_as(C.POINTER(C.c_int32)), None,
                           0, err, 256)
    if n < 0:
        raise OracleError(-n, err.value.decode())
    out = np.zeros(n)
    L.oracle_block_cgt(r, t.ctypes.data_as(C.POINTER(C.c_int32)), d.ctypes.data_as(C.POINTER(C.c_int32)),
                       out.ctypes.data_as(C.POINTER(C.c_double)), n, err, 256)
    shape = tuple(int(x) + 1 for x in tj)
    return out.reshape(shape, order="F") if r else out.reshape(())


def norm_ext(tj, dirs) -> float:
    t = _ivec(tj)
    d = _ivec(dirs)
    return lib().oracle_norm_ext(len(tj), t.ctypes.data_as(C.POINTER(C.c_int32)),
                                 d.ctypes.data_as(C.POINTER(C.c_int32)))


def x3(tja, dira, tjb, dirb, ia, ib) -> float:
    L = lib()
    P = C.POINTER
    a, da, b, db, ca, cb = (_ivec(v) for v in (tja, dira, tjb, dirb, ia, ib))
    out = C.c_double()
    err = C.create_string_buffer(256)
    rc = L.oracle_x3(len(tja), a.ctypes.data_as(P(C.c_int32)), da.ctypes.data_as(P(C.c_int32)), len(tjb),
                     b.ctypes.data_as(P(C.c_int32)), db.ctypes.data_as(P(C.c_int32)), ca.ctypes.data_as(P(C.c_int32)),
                     cb.ctypes.data_as(P(C.c_int32)), len(ia), C.byref(out), err, 256)
    if rc:
        raise OracleError(rc, err.value.decode())
    return out.value


# ---------------------------------------------------------------- block ops
# numpy restatements of permuteQS and plusQS (see the header's status list).

def _syms(T: QSpace):
    if not T.qtype:
        return ["A"] * T.qdim
    return [x.strip() for x in T.qtype.split(",")]


def _tag_dirs(tags, r):
    return [(-1 if t.endswith("*") else 1) for t in tags] if tags else [1] * r


def _block_labels(T: QSpace, i, j):
    return [int(T.Q[k][i, j]) for k in range(T.rank)]


def _weights(T: QSpace, i, dirs):
    """(NormCGW data factor, output weights) of block i: cgw / normExt per
    SU(2) symmetry (Source/QSpace.cc:3727-3805, CRef::normExt clebsch.cc:4738-4760)."""
    f, w_out = 1.0, []
    cgw = T.info.get("cgw")
    for j, s in enumerate(_syms(T)):
        if s != "SU2":
            w_out.append(1.0)
            continue
        ne = norm_ext(_block_labels(T, i, j), dirs)
        w = float(cgw[i, j]) if cgw is not None else ne
        f *= w / ne
        w_out.append(ne)
    return f, w_out


def _as_is(T: QSpace, fac=1.0) -> QSpace:
    dirs = _tag_dirs(T.itags, T.rank)
    cg = []
    for i in range(T.nblk):
        w = []
        for j, s in enumerate(_syms(T)):
            if s != "SU2":
                w.append(1.0 if T.info.get("cgw") is None else float(T.info["cgw"][i, j]))
            elif T.info.get("cgw") is not None:
                w.append(float(T.info["cgw"][i, j]))
            else:
                w.append(norm_ext(_block_labels(T, i, j), dirs))
        cg.append(w)
    data = [d.copy(order="F") if fac == 1.0 else fac * d for d in T.data]
    return QSpace([q.copy() for q in T.Q], data, T.qtype, list(T.itags),
                  np.array(cg).reshape(T.nblk, -1) if T.nblk else None)


def permute(A: QSpace, perm, conj=0) -> QSpace:
    """permuteQS(A, perm [,'conj']) -- Source/permuteQS.cc:25-48: a short
    permutation is completed with the identity (wbperm::Extend), a long one
    is an error; identity without conj returns A; else QSpace::permute
    (Source/QSpace.hh:2721-2780: labels, dims, itags and every block's data
    permuted; conj flips each leg's conj flag) and NormCGW
    (Source/QSpace.cc:3727-3805)."""
    r = A.rank
    perm = [int(x) for x in perm]
    if len(perm) > r:
        raise OracleError(1, f"invalid permutation (len={len(perm)}/{r})")
    p = perm + [i for i in range(r) if i not in perm]
    if sorted(p) != list(range(r)):
        raise OracleError(1, "invalid permutation")
    if r == 0 or (p == list(range(r)) and not conj):
        return _as_is(A)
    tags = [A.itags[k] for k in p] if A.itags else []
    if conj:
        tags = [t[:-1] if t.endswith("*") else t + "*" for t in tags]
    din, dout = _tag_dirs(A.itags, r), _tag_dirs(tags, r)
    Q = [A.Q[k].copy() for k in p]
    dconj = [-x for x in din] if conj else din
    data, cg = [], []
    for i, d in enumerate(A.data):
        f, _ = _weights(A, i, din)
        w = []
        for j, s in enumerate(_syms(A)):
            if s != "SU2":
                w.append(1.0)
                continue
            # the block's CGT is the canonical CGT of the new leg order: the
            # transposed CGT times the overlap sign of the two, which the
            # data take (no CGT permutation in the tensor format)
            tj0 = _block_labels(A, i, j)
            c_old = block_cgt(tj0, dconj)
            c_new = block_cgt([tj0[k] for k in p], [dconj[k] for k in p])
            f *= 1.0 if np.sum(c_new * np.transpose(c_old, p)) > 0 else -1.0
            w.append(norm_ext([int(Q[k][i, j]) for k in range(r)], dout))
        cg.append(w)
        x = np.transpose(np.asarray(d).reshape(d.shape[:r], order="F"), p)
        if conj:  # conj conjugates complex data (wbarray::permute, Source/wbarray.cc:2738)
            x = np.conj(x)
        data.append(np.asfortranarray(x * f))
    return QSpace(Q, data, A.qtype, tags, np.array(cg).reshape(A.nblk, -1) if A.nblk else None)


def _sorted_groups(R):
    """stable ascending lexicographic order of records (column 0 most
    significant, ties by index: Wb::hpsort / groupRecs, Source/wbsort.cc:88-143,
    Source/wbMatrix.cc:1682-1731) and the runs of equal records"""
    order = sorted(range(len(R)), key=lambda i: (tuple(R[i]), i))
    groups, cur = [], []
    for i in order:
        if cur and tuple(R[i]) != tuple(R[cur[0]]):
            groups.append(cur)
            cur = []
        cur.append(i)
    if cur:
        groups.append(cur)
    return groups


def _tags_equal(a, b):
    """iTags == with itag_::sameAs(.,'l') (Source/wbindex.cc:989-1007)"""
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if x.endswith("*") != y.endswith("*"):
            return False
        nx, ny = x.rstrip("*"), y.rstrip("*")
        if nx != ny and nx and ny:
            return False
    return True


def plus(A: QSpace, B: QSpace, bfac=1.0, skip_zero=1) -> QSpace:
    """plusQS(A, B [,bfac]) -- Source/plusQS.cc:17-55: C = A + bfac B.
    Abelian: QSpace::plus_plain (Source/QSpace.cc:3593-3705) over the sorted
    union of records (getQOverlapU, QSpace.cc:1266-1320).  SU(2) labels: Cat
    + MakeUnique (QSpace.cc:3075-3140,3807-3905): no common record -> A's
    blocks then B's; otherwise sorted unique records, zero blocks of a group
    not added.  Then SkipZeroData (QSpace.cc:3200-3270) and NormCGW."""
    if A.nblk == 0 and B.nblk == 0:
        return _as_is(A)
    both = A.nblk and B.nblk
    ref = A if A.nblk else B
    if both and (A.qdim != B.qdim or A.rank != B.rank):
        raise OracleError(3, "QSpace mismatch")
    if both and _syms(A) != _syms(B) and A.qtype and B.qtype:
        raise OracleError(3, "got inconsistent qtype")
    syms = _syms(ref)
    cg = "SU2" in syms
    r = ref.rank
    tags = list(B.itags) if (not A.itags or not A.nblk) else list(A.itags)
    if both and A.itags and B.itags:
        if cg:
            if not _tags_equal(A.itags, B.itags):
                raise OracleError(2, "itag inconsistency")
        else:
            if any(x.endswith("*") != y.endswith("*") for x, y in zip(A.itags, B.itags)):
                raise OracleError(2, "got inconsistent itags")
            if not _tags_equal(A.itags, B.itags):
                tags = ["*" if t.endswith("*") else "" for t in tags]
    RA = [tuple(A.QIDX[i]) for i in range(A.nblk)]
    RB = [tuple(B.QIDX[i]) for i in range(B.nblk)]
    groups = _sorted_groups(RA + RB)
    ga, gb, overlap = [], [], False
    for g in groups:
        if len(g) > 2:
            raise OracleError(4, "got non-unique QIDX")
        a = [i for i in g if i < A.nblk]
        b = [i - A.nblk for i in g if i >= A.nblk]
        if len(a) > 1 or len(b) > 1:
            raise OracleError(4, "non-unique QIDX")
        ga.append(a[0] if a else -1)
        gb.append(b[0] if b else -1)
        if a and b:
            overlap = True
            if A.data[a[0]].shape[:r] != B.data[b[0]].shape[:r]:
                raise OracleError(4, "block size mismatch")
    if cg and not overlap:
        oa = list(range(A.nblk)) + [-1] * B.nblk
        ob = [-1] * A.nblk + list(range(B.nblk))
    else:
        oa, ob = ga, gb
    dA, dB, dC = _tag_dirs(A.itags, r), _tag_dirs(B.itags, r), _tag_dirs(tags, r)
    Q, data, cgw = [[] for _ in range(r)], [], []
    for a, b in zip(oa, ob):
        if cg and a >= 0 and b >= 0 and overlap:
            za, zb = not np.any(A.data[a] != 0), not np.any(B.data[b] != 0)
            if za and not zb:
                a = -1
            elif zb:
                b = -1
        fa = (_weights(A, a, dA)[0] if cg else 1.0) if a >= 0 else 0.0
        fb = (bfac * _weights(B, b, dB)[0] if cg else bfac) if b >= 0 else 0.0
        if a >= 0 and b >= 0:
            x = fa * A.data[a]
            y = fb * B.data[b]
            d = x + y
        elif a >= 0:
            d = fa * A.data[a]
        else:
            d = fb * B.data[b]
        S, sb = (A, a) if a >= 0 else (B, b)
        for k in range(r):
            Q[k].append(S.Q[k][sb])
        data.append(np.asfortranarray(d))
        w = []
        for j, s in enumerate(syms):
            w.append(norm_ext(_block_labels(S, sb, j), dC) if s == "SU2" else 1.0)
        cgw.append(w)
    # SkipZeroData (eps = 0): exact-zero blocks dropped unless a rank-2
    # block-diagonal matrix
    q = ref.qdim
    blockdiag = r == 2 and all(np.array_equal(Q[0][i], Q[1][i]) for i in range(len(data)))
    keep = list(range(len(data)))
    if skip_zero and data and not (r == 2 and blockdiag):
        keep = [i for i in keep if np.any(data[i] != 0)]
    Qk = [np.array([Q[k][i] for i in keep], dtype=np.int32).reshape(len(keep), q) for k in range(r)]
    return QSpace(Qk, [data[i] for i in keep], A.qtype or B.qtype, tags,
                  np.array([cgw[i] for i in keep]).reshape(len(keep), -1) if keep else None)


# ------------------------------------------------------------- orthoQS (SVD)
def _mark_set_desc(E, Nkmin, Nkeep, Etrunc):
    """markSet(F,L,E,mark,Nkmin,Nkeep,Etrunc,b=0,dmax,'desc')
    (Source/gathered.cc:1957-2079) on E sorted ascending: marks the states
    kept from the top; returns (mark, Nkeep)"""
    N = len(E)
    mark = np.zeros(N, dtype=bool)
    if Nkeep < 0 or Nkeep > N:
        Nkeep = N
    if Nkeep == 0:
        return mark, 0
    if Nkeep >= N and Etrunc <= 0:
        mark[:] = True
        return mark, N
    if Etrunc > 0:
        lx = N - Nkmin if Nkmin < N else 0
        l = -1  # the unsigned loop running past index 0
        for i in range(N - 1, -1, -1):
            if i and E[i - 1] > E[i]:
                raise OracleError(1, "data not sorted")
            if E[i] <= Etrunc and i < lx:
                l = i
                break
        # (N-1-l) > Nkeep in unsigned arithmetic: l = -1 counts as N
        if (N if l < 0 else N - 1 - l) > Nkeep:
            l = N - 1 - Nkeep
    else:
        l = N - 1 - Nkeep
    if l < 0:
        mark[:] = True
        return mark, N
    if l == N - 1:
        return mark, 0
    mark[l + 1:] = True
    return mark, N - 1 - l


def _mark_set(SV, Nkmin, Nkeep, Etrunc):
    """markSet(E0,mark,Nkmin,Nkeep,Etrunc,E,eps=0,...,'desc')
    (Source/gathered.cc:1822-1884): all values sorted ascending (stable,
    Wb::hpsort), the inner markSet, marks mapped back per block"""
    E = np.concatenate(SV) if SV else np.zeros(0)
    N = len(E)
    marks = [np.zeros(len(s), dtype=bool) for s in SV]
    if N == 0:
        return marks, Nkeep
    if Nkeep < 0:
        Nkeep = N
    if Nkeep == 0:
        return marks, 0
    if Nkeep >= N and Etrunc <= 0:
        return [np.ones(len(s), dtype=bool) for s in SV], N
    P = np.argsort(E, kind="stable")
    mm, Nkeep = _mark_set_desc(E[P], Nkmin, Nkeep, Etrunc)
    m = np.zeros(N, dtype=bool)
    m[P] = mm
    o = 0
    for i, s in enumerate(SV):
        marks[i] = m[o:o + len(s)]
        o += len(s)
    return marks, Nkeep


def _tag_conj(t):
    return t[:-1] if t.endswith("*") else t + "*"


def ortho(PSI: QSpace, K, Nkmin=0, Nkeep=-1, stol=0.0, itag=""):
    """orthoQS(PSI,A1,A2,K,Nkmin,Nkeep,stol,tx) -- Source/mpsortho.cc:1031-1078:
    SVD_Data::blockSVD (:545-613) then dmrgTruncate (:615-954), for abelian
    symmetries, and for SU(2) (alone or with abelian factors) when r <= 3
    and K = r-1 or 1: the group matrix takes each block times prod cgw /
    sqrt(qdim_tot), the norms weight every singular value by qdim_tot; for
    K = r-1, A1 keeps PSI's CGT (data times sqrt(qdim_tot) / normExt,
    weight normExt) and A2 is the rank-2 identity (weight normExt =
    sqrt(d)); for K = 1 of a rank-3 tensor, A1 is the identity and A2
    carries PSI's CGT permuted to (legs 1, 2, new); info carries svd_qdim,
    the second column of the reference's svd records.
    Returns (A1, A2, info):
      * blocks grouped by the label sum over legs [0] (K=1), [r-1] (K=r-1) or
        [0..K-1] (QSpace::getQsum, Source/QSpace.hh:2385-2420), groups in
        stable ascending order (groupRecs);
      * per group the block matrix of QBlock::toBlockMatrix (:30-256): rows
        the sorted distinct label records of legs 0..K-1, columns those of
        legs K..r-1, each block in its column-major fused form;
      * economy SVD per group (wbSVD -> LAPACK gesvd, Source/wbarray_blas.cc:
        1090-1180), singular values descending;
      * markSet 'desc' over all singular values (Nkmin, Nkeep, stol);
        discarded weight s2t, norm readjusted by sfac = |S| / sqrt(|S|^2-s2t)
        when anything is discarded;
      * A1 = U[:, kept] per row record with legs (legs 0..K-1, new);
        A2 = (V S sfac)[:, kept] per column record with legs (legs K..r-1,
        new); itags tU / tV (:786-818); A2's blocks with norm <= stol*1e-14
        dropped unless A2 is a rank-2 block-diagonal matrix (SkipZeroData
        (xtol2,'b'), Source/QSpace.cc:3201-3270).
    info: svd (all singular values, descending), svd2tr, sv2tot, sfac,
    Nkeep, Ntot, flag (dmrgTruncate's count of notes)."""
    r, n = PSI.rank, PSI.nblk
    syms = _syms(PSI)
    su2 = [j for j, s in enumerate(syms) if s == "SU2"]
    if any(s not in ("A", "SU2") for s in syms):
        raise OracleError(5, f"symmetry {PSI.qtype} not restated")
    if su2 and (r > 3 or K not in (1, r - 1)):
        raise OracleError(5, "non-abelian symmetries need K = 1 or r-1 and rank <= 3 here (outer "
                             "multiplicities are not restated)")
    if PSI.is_complex():
        raise OracleError(5, "complex data")
    if n == 0:
        raise OracleError(1, "got empty PSI")
    if not (1 <= K < r):
        raise OracleError(1, f"index specs out of bounds ({K}/{r})")
    tags = list(PSI.itags)
    if len(tags) != r:
        raise OracleError(2, "got invalid/empty itags")
    dirs = _tag_dirs(tags, r)
    if K != 1:
        if any(dirs[i] <= 0 for i in range(K)):
            raise OracleError(2, "only ingoing indices can be fused into outgoing")
        if any(dirs[i] != dirs[K] for i in range(K + 1, r)):
            raise OracleError(2, "2nd/right group is mixed in+out")
    Ib = [0] if K == 1 else ([r - 1] if K == r - 1 else list(range(K)))
    Qb = [tuple(int(sum(PSI.Q[j][i, q] for j in Ib)) for q in range(PSI.qdim)) for i in range(n)]
    groups = _sorted_groups(Qb)
    QB = []
    for g in groups:
        rows = [tuple(int(x) for j in range(K) for x in PSI.Q[j][i]) for i in g]
        cols = [tuple(int(x) for j in range(K, r) for x in PSI.Q[j][i]) for i in g]
        g1 = _sorted_groups(rows)
        g2 = _sorted_groups(cols)
        ri = {}
        for a, idx in enumerate(g1):
            for t in idx:
                ri[t] = a
        ci = {}
        for b, idx in enumerate(g2):
            for t in idx:
                ci[t] = b
        S1 = [None] * len(g1)
        S2 = [None] * len(g2)
        seen = set()
        for t, i in enumerate(g):
            a, b = ri[t], ci[t]
            if (a, b) in seen:
                raise OracleError(4, "block not unique")
            seen.add((a, b))
            s = PSI.data[i].shape + (1,) * (r - PSI.data[i].ndim)
            s1, s2 = tuple(s[:K]), tuple(s[K:r])
            if S1[a] is None:
                S1[a] = s1
            elif S1[a] != s1:
                raise OracleError(4, "block size mismatch")
            if S2[b] is None:
                S2[b] = s2
            elif S2[b] != s2:
                raise OracleError(4, "block size mismatch")
        D1 = [int(np.prod(s)) for s in S1]
        D2 = [int(np.prod(s)) for s in S2]
        cD1 = np.concatenate([[0], np.cumsum(D1)]).astype(int)
        cD2 = np.concatenate([[0], np.cumsum(D2)]).astype(int)
        # multiplet dimension of the grouping leg (QBlock::toBlockMatrix,
        # Source/mpsortho.cc:62-90: leg 0 for K=1, leg K otherwise)
        l1 = 0 if K == 1 else K
        qd = 1
        for j in su2:
            qd *= int(PSI.Q[l1][g[0], j]) + 1
        MM = np.zeros((cD1[-1], cD2[-1]))
        for t, i in enumerate(g):
            a, b = ri[t], ci[t]
            # block scale cfac = prod cgw / sqrt(qdim_tot) (mpsortho.cc:192-210)
            cf = 1.0
            for j in su2:
                cf *= _cgw_of(PSI, i, j, dirs)
            cf /= np.sqrt(float(qd))
            blk = np.asarray(PSI.data[i]).reshape(D1[a], D2[b], order="F")
            MM[cD1[a]:cD1[a + 1], cD2[b]:cD2[b + 1]] = blk if cf == 1.0 else blk * cf
        U, S, Vt = np.linalg.svd(MM, full_matrices=False)
        QB.append(dict(g=g, rows=[rows[x[0]] for x in g1], cols=[cols[x[0]] for x in g2], S1=S1, S2=S2,
                       cD1=cD1, cD2=cD2, U=U, S=S, Vc=Vt.T, key=Qb[g[0]], qd=qd))
    SV = [b["S"] for b in QB]
    N = sum(len(s) for s in SV)
    if N == 0:
        raise OracleError(1, "got empty SVD data")
    # norms weighted by the multiplet dimension (dmrgTruncate :668-684,
    # QBlock::get_s2t Source/mpsortho.hh:98-106); the svd records are
    # (S, qdim) sorted descending (SortRecs(P,-1))
    snorm2 = sum(float(np.sum(b["S"] * b["S"])) * b["qd"] for b in QB)
    s_all = np.concatenate(SV)
    d_all = np.concatenate([np.full(len(b["S"]), float(b["qd"])) for b in QB])
    order = np.lexsort((-d_all, -s_all))
    SM, SMd = s_all[order], d_all[order]
    marks, NK = _mark_set(SV, max(Nkmin, 0), Nkeep, stol)
    s2t = sum(float(np.sum(b["S"][~m] ** 2)) * b["qd"] for b, m in zip(QB, marks))
    nk = sum(int(m.sum()) for m in marks)
    flag = 0
    xtol2 = stol * 1e-14 if stol > 0 else 0.0
    unlimited = Nkeep < 0
    if SM[-1] > stol:
        if not unlimited and len(SM) > Nkeep:
            if stol:
                flag += 1
        else:
            xtol2 = 0.0
    else:
        if not unlimited and NK < Nkeep:
            pass
        elif NK < len(SM):
            flag += 1
    sfac = 1.0
    if s2t:
        if snorm2 > s2t:
            sfac = np.sqrt(snorm2) / np.sqrt(snorm2 - s2t)
            if abs(sfac - 1) > 0.01:
                flag += 1
        elif snorm2 != s2t:
            raise OracleError(1, "discarded weight exceeds the norm")
    # itags of the new leg (dmrgTruncate, Source/mpsortho.cc:786-818)
    tx = itag.rstrip("*")
    if r - K == 1:
        last = tags[-1]
        tu = last if not tx else (tx + "*" if last.endswith("*") else tx)
        if not tu.endswith("*") and nk:
            is_op = r == 2 and tags[0].endswith("*") != tags[1].endswith("*")
            if not is_op:
                raise OracleError(5, "new leg incoming (the reference's identity-CG flip, revert_Uidx, "
                                     "is not restated)")
    elif K == 1:
        first = tags[0]
        tu = first if not tx else (tx + "*" if first.endswith("*") else tx)
        tu = _tag_conj(tu)
    else:
        tu = tx if tags[0].endswith("*") else tx + "*"
    tU = tags[:K] + [tu]
    tV = tags[K:] + [_tag_conj(tu)]
    QU, DU, QV, DV = [[] for _ in range(K + 1)], [], [[] for _ in range(r - K + 1)], []
    WU, WV = [], []
    dU, dV = _tag_dirs(tU, K + 1), _tag_dirs(tV, r - K + 1)
    q = PSI.qdim

    def weights(recs, dd):
        """output weights: normExt per SU(2) symmetry (NormSignW with
        useExt, Source/clebsch.cc:5035-5052; the identity CGT of a rank-2
        A2 has cgw = sqrt(d) = normExt, CRef::initIdentityR), 1 otherwise"""
        return [norm_ext([rr[j] for rr in recs], dd) if j in su2 else 1.0 for j in range(q)]
    for b, m in zip(QB, marks):
        Ik = np.nonzero(m)[0]
        if not len(Ik):
            continue
        if r - K == 1:
            new = b["cols"][0]
        else:
            left = b["rows"][0]
            new = tuple(sum(left[j * q + x] for j in range(K)) for x in range(q))
        for a, rec in enumerate(b["rows"]):
            if b["cD1"][a + 1] == b["cD1"][a]:
                continue  # empty sub-block: no block (initFromBlockMatrix)
            recs = [rec[j * q:(j + 1) * q] for j in range(K)] + [new]
            for j in range(K + 1):
                QU[j].append(recs[j])
            blk = b["U"][b["cD1"][a]:b["cD1"][a + 1], Ik]
            w = weights(recs, dU)
            if su2 and K == r - 1:
                # U keeps PSI's CGT: weight sqrt(qdim_tot) renormalised to
                # normExt (init_wId + NormSignW, mpsortho.cc:447-452)
                blk = blk * (np.sqrt(float(b["qd"])) / float(np.prod([w[j] for j in su2])))
            WU.append(w)
            DU.append(np.asfortranarray(blk.reshape(b["S1"][a] + (len(Ik),), order="F")))
        VS = b["Vc"][:, Ik] * (b["S"][Ik] * sfac)[None, :]
        XQ, XD = [], []
        for c, rec in enumerate(b["cols"]):
            if b["cD2"][c + 1] == b["cD2"][c]:
                continue
            XQ.append([rec[j * q:(j + 1) * q] for j in range(r - K)] + [new])
            blk = VS[b["cD2"][c]:b["cD2"][c + 1], :]
            XD.append(np.asfortranarray(blk.reshape(b["S2"][c] + (len(Ik),), order="F")))
        if su2 and K == 1 and r == 3:
            XQ, XD = _su2_first_leg_a2(PSI, dirs, b["qd"], XQ, XD, su2)
        blockdiag = r - K + 1 == 2 and all(tuple(x[0]) == tuple(x[1]) for x in XQ)
        for xq, xd in zip(XQ, XD):
            if not blockdiag and float(np.sum(xd * xd)) <= xtol2 * xtol2:
                continue
            for j in range(r - K + 1):
                QV[j].append(xq[j])
            WV.append(weights(xq, dV))
            DV.append(xd)

    def mk(Qs, D, W, tg):
        # an SU(2) result keeps its weight columns when nothing is kept
        cg = np.array(W, dtype=np.float64).reshape(len(D), q) if su2 else None
        return QSpace([np.array(x, dtype=np.int32).reshape(len(D), q) for x in Qs], D, PSI.qtype, tg, cg)

    info = dict(svd=SM, svd2tr=s2t / snorm2, sv2tot=snorm2, sfac=sfac, Nkeep=nk, Ntot=len(SM), flag=flag,
                Nkmin=max(Nkmin, 0), stol=stol)
    if su2:
        info["svd_qdim"] = SMd
    return mk(QU, DU, WU, tU), mk(QV, DV, WV, tV), info


def _su2_first_leg_a2(PSI: QSpace, dirs, qd, XQ, XD, su2):
    """A2 blocks of a non-abelian K = 1 split of a rank-3 tensor
    (initFromBlockMatrix with cgp2 = [1,2,0], Source/mpsortho.cc:440-452):
    PSI's CGT with weight sqrt(qdim_tot) renormalised to normExt on legs
    (new, 1, 2) -- the labels and directions of PSI's block -- then that CGT
    permuted to (1, 2, new), restated as permute() of that tensor."""
    q, n = PSI.qdim, len(XD)
    Qx = [[xq[2] for xq in XQ], [xq[0] for xq in XQ], [xq[1] for xq in XQ]]
    data = []
    for xq, xd in zip(XQ, XD):
        ne = 1.0
        for j in su2:
            ne *= norm_ext([int(xq[2][j]), int(xq[0][j]), int(xq[1][j])], dirs)
        data.append(np.asfortranarray(np.transpose(xd, (2, 0, 1)) * (np.sqrt(float(qd)) / ne)))
    X = QSpace([np.array(x, dtype=np.int32).reshape(n, q) for x in Qx], data, PSI.qtype, list(PSI.itags))
    Y = permute(X, [1, 2, 0])
    recs = [[tuple(int(v) for v in Y.Q[k][i]) for k in range(3)] for i in range(n)]
    assert all(tuple(a) == tuple(b) for rec, xq in zip(recs, XQ) for a, b in zip(rec, xq))
    return recs, [np.asfortranarray(d) for d in Y.data]


def _cgw_of(T: QSpace, i, j, dirs) -> float:
    """CGT weight of block i for symmetry j (normExt when T carries none)"""
    cgw = T.info.get("cgw")
    if cgw is not None:
        return float(cgw[i, j])
    return norm_ext(_block_labels(T, i, j), dirs)


def find_openblas():
    """numpy's bundled ILP64 OpenBLAS (path, dgemm symbol, threads symbol)."""
    import glob
    import numpy
    base = os.path.dirname(os.path.dirname(numpy.__file__))
    cands = glob.glob(os.path.join(base, "numpy.libs", "libscipy_openblas64_*.so")) + \
        glob.glob(os.path.join(base, "scipy_openblas64", "lib", "libscipy_openblas64_*.so"))
    for c in cands:
        return c, "scipy_dgemm_64_", "scipy_openblas_set_num_threads64_"
    return None


class CpuBaseline:
    """Reference-structure CPU data half (OpenMP over output blocks, per-pair
    host dgemm + axpy); structure pass and X-symbols done at construction."""

    def __init__(self, A: QSpace, ia, B: QSpace, ib, conjA=0, conjB=0):
        L = lib()
        ob = find_openblas()
        if ob is not None:
            L.oracle_load_blas(ob[0].encode(), ob[1].encode(), ob[2].encode())
        ca, self._ka = to_c(A)
        cb, self._kb = to_c(B)
        ia_ = _ivec(ia)
        ib_ = _ivec(ib)
        err = C.create_string_buffer(512)
        self.h = L.oracle_baseline_create(C.byref(ca), ia_.ctypes.data_as(C.POINTER(C.c_int32)), len(ia), conjA,
                                          C.byref(cb), ib_.ctypes.data_as(C.POINTER(C.c_int32)), len(ib), conjB,
                                          err, 512)
        if not self.h:
            raise OracleError(1, err.value.decode())
        self.flops = L.oracle_baseline_flops(self.h)
        self.blas = L.oracle_blas_name().decode()

    def run(self, nthreads):
        return lib().oracle_baseline_run(self.h, nthreads)

    def output(self):
        L = lib()
        n = L.oracle_baseline_output(self.h, None, 0)
        out = np.zeros(max(1, n))
        L.oracle_baseline_output(self.h, out.ctypes.data_as(C.POINTER(C.c_double)), n)
        return out[:n]

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_baseline_destroy(self.h)
            self.h = None
