"""RC_STORE CData files of SU(2) Clebsch-Gordan tensors (read, and write in
the same layout).

The reference keeps the CGTs of its non-abelian symmetries on disk
(RCStore::save_CData / load_CData, Source/clebsch_io.cc:313-441; the store
class in Source/clebsch.hh, RCStore):

  * path: ``<RC_STORE>/<type>/CStore/<qdir>/(<labels>).cgd``
    (RCStore::get_file_name for a QSet, Source/clebsch_io.cc:192-226, and
    get_rcs_path :20-190; system/RCStore.jl qset2fname).  ``qdir`` is the
    direction string ('+' incoming, '-' outgoing) of the legs in the
    CData's sorted order (QSet::Sort, ``canonical_legs``); ``labels`` are
    the legs' labels in that order, each in compact form
    (one character per label, 0-9 then A-Z for 10-35, CG::qset2cstr) when
    every label is below 36, else decimal, legs separated by ',' and by
    ';' where the direction changes (QSet::QStrS, Source/clebsch.cc:1669-1710);
  * contents: a MATLAB file with ``CRef``, the CData's header struct
    (CData::add2MxStruct, clebsch_io.cc:1343-1370: type, qset, qdir, cid
    and an empty cgd), and ``cdata``, the CGT as a sparse array
    (wbsparray::toMx, Source/wbsparray.hh:1144-1154): rank <= 2 as a MATLAB
    sparse matrix, otherwise a struct {S, data} with S the dimensions and
    data a sparse d1 x d2 matrix over the column-major linear index
    (wbsparray::add2MxStruct / get2DIndex, Source/wbsparray.cc:2026-2040,
    Source/wbsparray.hh:1838-1890).  A CGT with outer multiplicity has one
    dimension more than its rank (S of length r + 1; system/RCStore.jl
    getOM), the basis index last;
  * cid: cgdStatus::toMx (clebsch_io.cc:1194-1203), [ctime, mtime, ID,
    type, type flags, rtype].

This module reads such files into dense arrays and relates them to this
library's bases (``qsx_su2_om_basis``): ``basis_overlap`` gives the matrix
that carries coefficients on the stored basis to this library's, which is
what data written with the reference's outer-multiplicity bases need before
they enter a contraction here.  No RC_STORE written by the reference exists
in this repository, so the reader is tested on files written in the layout
above (parity with real stores is unpinned).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

__all__ = ["CData", "store_root", "label_string", "cdata_path", "read_cdata", "write_cdata", "basis_overlap",
           "find_cdata", "canonical_legs"]


def store_root() -> str | None:
    """the store directory of the environment (RC_STORE, as RCStore::setupDirs)"""
    return os.environ.get("RC_STORE") or None


def canonical_legs(tj, dirs):
    """block legs in the stored (sorted QSet) order: QSet::Sort,
    Source/clebsch.cc:1435-1505, as the library applies it to its CGTs"""
    from .api import su2_canon_order
    return su2_canon_order(tj, dirs)


def _compact(q: int) -> str:
    return chr(ord("0") + q) if q < 10 else chr(ord("A") + q - 10)


def label_string(labels, dirs) -> str:
    """'(1,1;2)' for labels (1, 1, 2) with directions (+, +, -)"""
    labels = [int(x) for x in labels]
    compact = all(0 <= q < 36 for q in labels)
    out = []
    for i, q in enumerate(labels):
        out.append(_compact(q) if compact else str(q))
        if i + 1 < len(labels):
            out.append(";" if dirs[i + 1] != dirs[i] else ",")
    return "(" + "".join(out) + ")"


def cdata_path(root: str, qtype: str, labels, dirs) -> str:
    """file of the CData with the given (sorted) labels and directions"""
    qdir = "".join("+" if d > 0 else "-" for d in dirs)
    return os.path.join(root, qtype, "CStore", qdir, label_string(labels, dirs) + ".cgd")


@dataclass
class CData:
    qtype: str
    qset: np.ndarray              # labels per leg (2S for SU(2))
    qdir: str                     # '+' / '-' per leg
    cid: np.ndarray = field(default_factory=lambda: np.zeros(6))
    cgt: np.ndarray = None        # dense, dims (2S+1 per leg [, om])

    @property
    def rank(self) -> int:
        return len(self.qdir)

    @property
    def om(self) -> int:
        return int(self.cgt.shape[self.rank]) if self.cgt is not None and self.cgt.ndim > self.rank else 1

    def basis(self) -> np.ndarray:
        """the stored basis tensors, array (om, dims...)"""
        c = self.cgt
        if c.ndim > self.rank:
            return np.moveaxis(c, self.rank, 0)
        return c[None]


def _field(s, name):
    v = s[name]
    while isinstance(v, np.ndarray) and v.dtype == object and v.size == 1:
        v = v.ravel()[0]
    return v


def _str(v) -> str:
    if isinstance(v, str):
        return v
    v = np.asarray(v)
    if v.size == 0:
        return ""
    if v.dtype.kind in "US":
        return "".join(str(x) for x in v.ravel(order="F"))
    return ""


def _dense_cgt(c, rank):
    import scipy.sparse as sp
    if sp.issparse(c):
        return np.asarray(c.toarray(), dtype=np.float64).reshape(c.shape if rank >= 2 else (-1,), order="F")
    if isinstance(c, np.ndarray) and c.dtype.names and "S" in c.dtype.names:
        S = [int(x) for x in np.ravel(_field(c, "S"))]
        data = _field(c, "data")
        if sp.issparse(data):
            flat = np.asarray(data.toarray(), dtype=np.float64).ravel(order="F")
        else:
            flat = np.asarray(data, dtype=np.float64).ravel(order="F")
        return flat.reshape(S, order="F")
    return np.asarray(c, dtype=np.float64)


def read_cdata(path: str) -> CData:
    """one .cgd file -> CData (dense CGT in the stored leg order)"""
    from scipy.io import loadmat
    m = loadmat(path, squeeze_me=False, struct_as_record=True, mat_dtype=True)
    if "CRef" not in m or "cdata" not in m:
        raise ValueError(f"{path}: not a CData file (needs CRef and cdata)")
    R = m["CRef"]
    qtype = _str(_field(R, "type"))
    qset = np.ravel(np.asarray(_field(R, "qset"), dtype=np.float64)).astype(np.int64)
    qdir = _str(_field(R, "qdir"))
    cid = np.ravel(np.asarray(_field(R, "cid"), dtype=np.float64)) if "cid" in R.dtype.names else np.zeros(6)
    r = len(qdir)
    cgt = _dense_cgt(m["cdata"], r)
    if cgt.ndim < r:
        cgt = cgt.reshape(cgt.shape + (1,) * (r - cgt.ndim), order="F")
    want = tuple(int(q) + 1 for q in qset) if qtype == "SU2" else None
    if want is not None and tuple(cgt.shape[:r]) != want:
        raise ValueError(f"{path}: CGT dims {cgt.shape} do not match the labels {tuple(qset)}")
    if cgt.ndim > r + 1:
        raise ValueError(f"{path}: CGT of rank {cgt.ndim} for {r} legs")
    return CData(qtype, qset, qdir, cid, cgt)


def write_cdata(path: str, cd: CData) -> None:
    """write a CData in the store layout (header struct + sparse data)"""
    import scipy.sparse as sp
    from scipy.io import savemat
    os.makedirs(os.path.dirname(path), exist_ok=True)
    c = np.asarray(cd.cgt, dtype=np.float64)
    hdr = {"type": cd.qtype, "qset": np.asarray(cd.qset, dtype=np.float64).reshape(1, -1), "qdir": cd.qdir,
           "cid": np.asarray(cd.cid, dtype=np.float64).reshape(1, -1), "cgd": np.zeros((0, 0))}
    if c.ndim <= 2:
        data = sp.csc_matrix(c.reshape(c.shape + (1,) * (2 - c.ndim)))
    else:
        # rows: the first m dims with prod >= sqrt(numel) (get2DSize)
        S = list(c.shape)
        d, m = S[0], 1
        while m < len(S) and d < np.sqrt(c.size):
            d *= S[m]
            m += 1
        d1 = int(np.prod(S[:m]))
        data = {"S": np.asarray(S, dtype=np.float64).reshape(1, -1),
                "data": sp.csc_matrix(c.reshape(d1, -1, order="F"))}
    savemat(path, {"CRef": hdr, "cdata": data}, do_compression=False, oned_as="row")


def find_cdata(root: str, qtype: str, labels, dirs) -> CData | None:
    """the stored CData of a sorted label set, or None if the store lacks it"""
    p = cdata_path(root, qtype, labels, dirs)
    return read_cdata(p) if os.path.exists(p) else None


def basis_overlap(cd: CData, ours: np.ndarray) -> np.ndarray:
    """T[a, b] = <stored_a | ours_b> over the legs in the stored order.
    This library's basis is orthonormal and spans the labels' invariant
    space, so coefficients x_a on the stored basis (any basis of that space)
    become y_b = sum_a x_a T[a, b] on this library's, exactly."""
    S = cd.basis()
    if S.shape[1:] != ours.shape[1:]:
        raise ValueError(f"basis shapes differ: {S.shape} vs {ours.shape}")
    return S.reshape(len(S), -1) @ ours.reshape(len(ours), -1).T
