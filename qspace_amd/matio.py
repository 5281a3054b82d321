"""QSpace tensors in MATLAB .mat files (the reference's struct {Q, data, info}).

Layout restated from the reference's MATLAB interface:
  * ``Q``: 1 x r cell, one n x QDIM double matrix of labels per leg
    (QSpace::QIDX_toMx, Source/QSpace.hh:2976-2995; read back by
    QSpace::init, :3150-3230, which concatenates the cells into QIDX and
    requires unique rows);
  * ``data``: n x 1 cell of column-major double arrays; MATLAB drops
    trailing singleton dimensions, which the reader appends again
    (appendSingletons, :3340-3344);
  * ``info``: 1 x 1 struct with fields qtype, otype, itags, fdir, ctime,
    cgr (INFO_toMx, :2998-3015); itags is a cell of strings with a trailing
    '*' for outgoing legs; cgr is an n x nsym struct array of CRef records
    with fields type, qset, qdir, cid, size, nnz, cgw, cgt
    (CRef::mxCreateStruct / add2MxStruct, Source/clebsch_io.cc:1722-1768).

Read: labels, data, itags, qtype and, per block and non-abelian symmetry,
the scalar weight cgw (outer multiplicity 1).  The CGT itself is not read:
this library's block CGT is the canonical CGT of the block's labels and
directions in its leg order.  The reference writes only CRefs in that
standard form: add2MxStruct refuses a CRef whose CGT permutation is not the
sorting permutation of its labels (isSortedDegQ, Source/clebsch_io.cc:1748-1756,
Source/clebsch.cc:4804-4824), and writes qset / qdir in the tensor's leg
order (the CData's labels adapted by the permutation).  A CRef whose legs
were permuted is brought to that form by CRef::Permute -> SortDegQ ->
X3Map::contractDegQ, which moves the sign of the transposition into cgw
(Source/clebsch.cc:5005-5031,4827-4868; Source/clebsch_aux.cc:4512-4600).
The reader checks each record's qset / qdir against the block's labels and
leg directions and raises ValueError on a mismatch (a CGT in another gauge,
which this format cannot represent).
Outer multiplicity: a block with one data dimension more than the rank
holds coefficients on a basis of its CGT space; this library reads and
writes them on its own basis (files marked info.otype = OM_OTYPE) and
refuses them in other files, where they refer to the reference's RC_STORE
basis (rcstore.basis_overlap converts them).
Write: the same struct with cgr records holding type, qset (the block's
labels of that symmetry in leg order), qdir, size, nnz and cgw; cid and cgt
are left empty (the reference recomputes CGTs from its own store).
"""
from __future__ import annotations

import numpy as np

from .qspace import QSpace

__all__ = ["qspace_from_mx", "qspace_to_mx", "load_qspace", "save_qspace", "OM_OTYPE"]

# info.otype of files whose blocks carry outer-multiplicity coefficients on
# this library's basis (qsx_su2_om_basis); such a trailing data dimension in
# any other file holds coefficients on the reference's own (RC_STORE) basis
OM_OTYPE = "qsx-om"


def _cell_list(a):
    """elements of a MATLAB cell (numpy object array) in column-major order"""
    a = np.asarray(a, dtype=object)
    return list(a.ravel(order="F"))


def _str(a) -> str:
    if isinstance(a, str):
        return a
    a = np.asarray(a)
    if a.size == 0:
        return ""
    if a.dtype.kind in "US":
        return "".join(str(x) for x in a.ravel(order="F"))
    if a.dtype == object:
        return _str(a.ravel(order="F")[0])
    return ""


def _field(s, name):
    """field of a 1 x 1 MATLAB struct loaded by scipy (structured array or
    dict); None if absent"""
    if isinstance(s, dict):
        return s.get(name)
    s = np.asarray(s)
    if s.dtype.names is None or name not in s.dtype.names:
        return None
    return s.ravel(order="F")[0][name]


def qspace_from_mx(S) -> QSpace:
    """QSpace from a MATLAB struct {Q, data, info} as scipy.io.loadmat
    returns it (struct_as_record=True, squeeze_me=False) or as a dict."""
    aQ, aD, aI = _field(S, "Q"), _field(S, "data"), _field(S, "info")
    if aQ is None or aD is None:
        raise ValueError("invalid QSpace (Q or data missing)")
    qcells = _cell_list(aQ)
    r = len(qcells)
    Q = [np.asarray(q, dtype=np.float64) for q in qcells]
    n = Q[0].shape[0] if r else 0
    qdim = Q[0].shape[1] if r and Q[0].ndim == 2 else 1
    for q in Q:
        if q.shape[0] != n:
            raise ValueError("invalid QSpace: size mismatch in Q data")
    Qi = []
    for q in Q:
        qi = np.rint(q).astype(np.int32).reshape(n, qdim)
        if not np.array_equal(qi, q.reshape(n, qdim)):
            raise ValueError("QSpace labels must be integers")
        Qi.append(qi)
    dcells = _cell_list(aD)
    if len(dcells) != n:
        raise ValueError(f"dimension mismatch (QIDX: {n} records, data: {len(dcells)})")
    ot = _field(aI, "otype") if aI is not None and np.asarray(aI).size else None
    own_om = ot is not None and _str(ot) == OM_OTYPE
    data = []
    for d in dcells:
        a = np.asarray(d)
        a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64, copy=False)
        if a.ndim > r + 1 or (a.ndim == r + 1 and not own_om):
            raise ValueError(f"data block of rank {a.ndim} > {r}: outer-multiplicity coefficients on the "
                             f"reference's RC_STORE basis (convert them with rcstore.basis_overlap)")
        if a.ndim <= r:
            a = a.reshape(a.shape + (1,) * (r - a.ndim), order="F")
        data.append(np.asfortranarray(a))
    qtype, itags, cgw = "", [], None
    if aI is not None and np.asarray(aI).size:
        qtype = _str(_field(aI, "qtype") if _field(aI, "qtype") is not None else "")
        it = _field(aI, "itags")
        if it is not None and np.asarray(it).size:
            itags = [_str(x) for x in _cell_list(it)]
            if len(itags) != r:
                raise ValueError(f"invalid number of itags ({len(itags)}/{r})")
        cg = _field(aI, "cgr")
        if cg is not None and np.asarray(cg).size:
            cg = np.asarray(cg)
            syms = [x.strip() for x in qtype.split(",")] if qtype else ["A"] * qdim
            if cg.ndim == 2 and cg.shape[0] == n:
                recs = cg
            elif cg.size == n * len(syms):
                recs = cg.reshape((n, len(syms)), order="F")
            else:
                raise ValueError(f"info.cgr of {cg.size} records for {n} blocks x {len(syms)} symmetries")
            cgw = np.ones((n, len(syms)))
            dirs = [(-1 if t.endswith("*") else 1) for t in itags] if itags else None
            for i in range(n):
                for j, sy in enumerate(syms):
                    if sy == "SU2" and j < recs.shape[1]:
                        _check_cgr_record(recs[i, j], [int(Qi[k][i, j]) for k in range(r)], dirs, i, j)
                        w = np.asarray(recs[i, j]["cgw"], dtype=np.float64).ravel()
                        if w.size > 1:
                            raise ValueError("outer multiplicity > 1 (cgw with several entries) not supported")
                        if w.size == 1:
                            cgw[i, j] = float(w[0])
            if "SU2" not in syms:
                cgw = None
    return QSpace(Qi, data, qtype, itags, cgw)


def _record_field(rec, name):
    names = getattr(getattr(rec, "dtype", None), "names", None)
    if names is not None:
        return rec[name] if name in names else None
    if isinstance(rec, dict):
        return rec.get(name)
    return None


def _check_cgr_record(rec, labels, dirs, i, j):
    """a cgr record's qset / qdir must be the block's labels of that symmetry
    and the legs' directions, in leg order (the standard CRef form the
    reference writes, Source/clebsch_io.cc:1748-1768); an empty field is not
    checked"""
    qs = _record_field(rec, "qset")
    if qs is not None and np.asarray(qs).size:
        q = np.asarray(qs, dtype=np.float64).ravel()
        if q.size != len(labels) or not np.array_equal(q, np.asarray(labels, dtype=np.float64)):
            raise ValueError(f"info.cgr({i + 1},{j + 1}).qset {q.tolist()} differs from the block's labels "
                             f"{labels} (a CGT in a non-standard gauge)")
    qd = _record_field(rec, "qdir")
    if qd is not None and dirs is not None:
        d = _str(qd)
        if d:
            want = "".join("+" if x > 0 else "-" for x in dirs)
            if d != want:
                raise ValueError(f"info.cgr({i + 1},{j + 1}).qdir '{d}' differs from the leg directions '{want}'")


def _norm_ext(tj, dirs) -> float:
    """external norm of a block CGT (CRef::normExt, Source/clebsch.cc:4738-4760):
    sqrt of the dimension of the leg whose direction is in the minority for
    rank 3, sqrt of the multiplet dimension for rank 2, 1 for scalars"""
    r = len(tj)
    if r == 0 or r > 3 or all(t == 0 for t in tj):
        return 1.0
    if r == 3:
        nin = sum(1 for d in dirs if d > 0)
        if nin in (0, 3):
            return 1.0
        odd = [k for k in range(3) if (nin == 2 and dirs[k] < 0) or (nin == 1 and dirs[k] > 0)][-1]
        return float(np.sqrt(tj[odd] + 1))
    return float(np.sqrt(max(tj) + 1))


def qspace_to_mx(T: QSpace) -> dict:
    """the MATLAB struct {Q, data, info} of T as a dict for scipy.io.savemat"""
    r, n = T.rank, T.nblk
    Qc = np.empty((1, r), dtype=object)
    for k in range(r):
        Qc[0, k] = np.asarray(T.Q[k], dtype=np.float64)
    Dc = np.empty((n, 1 if n else 0), dtype=object)
    for i in range(n):
        Dc[i, 0] = np.asfortranarray(T.data[i])
    itags = np.empty((1, len(T.itags)), dtype=object)
    for k, t in enumerate(T.itags):
        itags[0, k] = t
    syms = [x.strip() for x in T.qtype.split(",")] if T.qtype else []
    has_om = n and r and int(max(T.om())) > 1
    info = {"qtype": T.qtype, "otype": OM_OTYPE if has_om else "", "itags": itags, "fdir": np.zeros((0, 0)),
            "ctime": 0.0}
    if "SU2" in syms and n:
        dirs = [(-1 if t.endswith("*") else 1) for t in T.itags] if T.itags else [1] * r
        fields = ["type", "qset", "qdir", "cid", "size", "nnz", "cgw", "cgt"]
        cgr = np.zeros((n, len(syms)), dtype=[(f, object) for f in fields])
        for i in range(n):
            for j, sy in enumerate(syms):
                tj = [int(T.Q[k][i, j]) for k in range(r)]
                cgr[i, j]["type"] = sy
                cgr[i, j]["qset"] = np.asarray(tj, dtype=np.float64).reshape(1, -1)
                cgr[i, j]["qdir"] = "".join("+" if d > 0 else "-" for d in dirs)
                cgr[i, j]["cid"] = np.zeros((0, 0))
                cgr[i, j]["cgt"] = np.zeros((0, 0))
                if sy == "SU2":
                    cgr[i, j]["size"] = np.asarray([t + 1 for t in tj], dtype=np.float64).reshape(1, -1)
                    w = T.info["cgw"][i, j] if T.info.get("cgw") is not None else _norm_ext(tj, dirs)
                    cgr[i, j]["cgw"] = float(w)
                else:
                    cgr[i, j]["size"] = np.ones((1, r))
                    cgr[i, j]["cgw"] = np.zeros((0, 0))
                cgr[i, j]["nnz"] = np.zeros((0, 0))
        info["cgr"] = cgr
    else:
        info["cgr"] = np.zeros((0, 0))
    return {"Q": Qc, "data": Dc, "info": info}


def load_qspace(path: str, name: str) -> QSpace:
    """the QSpace variable ``name`` of a MATLAB (v5) .mat file"""
    import scipy.io
    m = scipy.io.loadmat(path, struct_as_record=True, squeeze_me=False)
    if name not in m:
        raise KeyError(f"no variable '{name}' in {path}")
    return qspace_from_mx(m[name])


def save_qspace(path: str, tensors: dict) -> None:
    """write QSpaces {name: QSpace} to a MATLAB (v5) .mat file"""
    import scipy.io
    scipy.io.savemat(path, {k: qspace_to_mx(v) for k, v in tensors.items()}, do_compression=False)
