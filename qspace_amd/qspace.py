"""QSpace tensor layout and the ctypes marshalling of the qsx C-ABI structs.

A QSpace is the reference's struct {Q, data, info} (Source/QSpace.hh:1455-1475,
mx layout Source/QSpace.hh:3150-3364):

* ``Q``     list of ``rank`` int32 arrays, each ``nblk x qdim`` (MATLAB's 1 x r
            cell of label matrices; SU(2) labels are 2S);
* ``data``  list of ``nblk`` column-major float64 blocks with ``rank`` dims
            (a block with outer multiplicity om > 1 has a trailing data
            dimension of size om: its coefficients on the om basis tensors of
            ``su2_om_basis``, the extra data dimension of QS_FULL_OM blocks);
* ``info``  dict with ``qtype`` ("A", "SU2", "A,SU2", ...), ``itags`` (list of
            leg tags, trailing '*' = outgoing leg, Source/wbindex.hh:632-700)
            and ``cgw`` (nblk x nsym outer-multiplicity weights, or None for
            the external normalisation sqrt(d) of CRef::normExt,
            Source/clebsch.cc:4738-4760).
"""
from __future__ import annotations

import ctypes as C
import re

import numpy as np

__all__ = [
    "QSpace", "QsxTensor", "QsxTensorOut", "QsxCell", "QsxOpts", "pack_itag", "unpack_itag",
    "parse_ctridx", "qtype_syms", "to_c", "from_out",
]


class QsxTensor(C.Structure):
    _fields_ = [
        ("nblk", C.c_int32), ("rank", C.c_int32), ("qdim", C.c_int32),
        ("Q", C.POINTER(C.c_int32)), ("dims", C.POINTER(C.c_int64)),
        ("data", C.POINTER(C.POINTER(C.c_double))), ("itags", C.POINTER(C.c_uint64)),
        ("qtype", C.c_char_p), ("cgw", C.POINTER(C.c_double)),
        ("data_im", C.POINTER(C.POINTER(C.c_double))), ("om", C.POINTER(C.c_int32)),
    ]


class QsxTensorOut(C.Structure):
    _fields_ = [
        ("nblk", C.c_int32), ("rank", C.c_int32), ("qdim", C.c_int32), ("nsym", C.c_int32),
        ("Q", C.POINTER(C.c_int32)), ("dims", C.POINTER(C.c_int64)), ("offs", C.POINTER(C.c_int64)),
        ("data", C.POINTER(C.c_double)), ("itags", C.POINTER(C.c_uint64)), ("cgw", C.POINTER(C.c_double)),
        ("qtype", C.c_char * 32), ("npairs", C.c_int64), ("useful_flops", C.c_double),
        ("gpu_ms", C.c_double), ("data_im", C.POINTER(C.c_double)), ("om", C.POINTER(C.c_int32)),
    ]


class QsxCell(C.Structure):
    """one entry of the cell form (qsx_cell): a QSpace or a nested cell and
    the options after it"""


QsxCell._fields_ = [("T", C.POINTER(QsxTensor)), ("nsub", C.c_int32), ("sub", C.POINTER(QsxCell)),
                    ("nopt", C.c_int32), ("opt", C.POINTER(C.c_char_p)), ("nidx", C.c_int32),
                    ("idx", C.POINTER(C.c_int32))]


class QsxOpts(C.Structure):
    _fields_ = [
        ("mode", C.c_int32), ("verbose", C.c_int32), ("nslices", C.c_int32), ("skip_zero", C.c_int32),
        ("hconj", C.c_int32), ("fermA", C.POINTER(C.c_double)), ("fermB", C.POINTER(C.c_double)),
    ]


def pack_itag(tag: str) -> int:
    """'L*' -> uint64 with the 8 chars little-endian and bit 7 of byte 0 set
    for an outgoing leg (itag_::Conj, Source/wbindex.hh:632-700)."""
    t = tag or ""
    conj = 0
    while t.endswith("*"):
        conj ^= 1
        t = t[:-1]
    b = t.encode()[:8]
    v = int.from_bytes(b.ljust(8, b"\0"), "little")
    if conj:
        v ^= 0x80
    return v


def unpack_itag(v: int) -> str:
    conj = bool(v & 0x80)
    b = (v & ~0x80).to_bytes(8, "little").rstrip(b"\0")
    return b.decode(errors="replace") + ("*" if conj else "")


def qtype_syms(qtype: str, qdim: int) -> list[str]:
    if not qtype:
        return ["A"] * qdim
    s = [x.strip() for x in qtype.split(",")]
    return ["A" if x in ("A", "U1") else x for x in s]


def parse_ctridx(s) -> tuple[list[int], int]:
    """ctr-idx -> (0-based indices, conj flag).  Accepts 1-based numeric
    sequences or strings like '13', '1,3', '[1 3]', '13*' (ctrIdx::init,
    Source/wbindex.cc:482-530 and Wb::Str2Idx, Source/wbio.c:182-297)."""
    if isinstance(s, str):
        t = s.rstrip()
        conj = 0
        while t.endswith("*"):
            conj ^= 1
            t = t[:-1].rstrip()
        t = t.rstrip(",; ")
        tok = [x for x in re.split(r"[\s,;\[\]]+", t) if x]
        if len(tok) == 1 and re.fullmatch(r"[0-9A-Za-z]+", tok[0]):
            # compact form: one alnum token, each char an index (0-9, A-Z=10.., a-z=36..)
            idx = []
            for c in tok[0]:
                if c.isdigit():
                    idx.append(int(c))
                elif "A" <= c <= "Z":
                    idx.append(ord(c) - ord("A") + 10)
                else:
                    idx.append(ord(c) - ord("a") + 36)
        else:
            idx = [int(x, 0) for x in tok]
        if not idx or min(idx) < 1:
            raise ValueError(f"invalid contraction index '{s}' (1-based)")
        if len(set(idx)) != len(idx):
            raise ValueError(f"invalid index set '{s}' (not unique)")
        return [i - 1 for i in idx], conj
    idx = [int(x) for x in np.atleast_1d(s)]
    if not idx or min(idx) < 1:
        raise ValueError("contraction index must be 1-based")
    if len(set(idx)) != len(idx):
        raise ValueError("contraction index not unique")
    return [i - 1 for i in idx], 0


class QSpace:
    """Block-sparse symmetric tensor, reference struct layout."""

    def __init__(self, Q, data, qtype: str = "", itags=None, cgw=None):
        if isinstance(Q, np.ndarray) and Q.ndim == 2:
            raise TypeError("Q must be a list of per-leg label matrices")
        self.Q = [np.ascontiguousarray(np.asarray(q, dtype=np.int32)) for q in Q]
        n = self.Q[0].shape[0] if self.Q else 0
        self.Q = [q.reshape(n, -1) if n else q.reshape(0, q.shape[-1] if q.ndim == 2 else 1) for q in self.Q]
        r = len(self.Q)
        self.data = []
        for d in data:
            a = np.asarray(d)
            a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64, copy=False)
            if a.ndim < r:
                a = a.reshape(a.shape + (1,) * (r - a.ndim), order="F")
            self.data.append(np.asfortranarray(a))
        if cgw is not None:
            cgw = np.asarray(cgw, dtype=np.float64)
            # (n, nsym); an empty tensor keeps the column count it was given
            cgw = cgw.reshape(n, -1) if n else cgw.reshape(0, cgw.shape[-1] if cgw.ndim == 2 else 0)
        self.info = {"qtype": qtype, "itags": list(itags) if itags is not None else [], "cgw": cgw}
        if len(self.data) != n:
            raise ValueError(f"QSpace: {n} label records but {len(self.data)} data blocks")

    # -- shape -------------------------------------------------------------
    @property
    def rank(self) -> int:
        return len(self.Q)

    @property
    def nblk(self) -> int:
        return len(self.data)

    @property
    def qdim(self) -> int:
        return self.Q[0].shape[1] if self.Q else 1

    @property
    def qtype(self) -> str:
        return self.info["qtype"]

    @property
    def itags(self) -> list[str]:
        return self.info["itags"]

    @property
    def QIDX(self) -> np.ndarray:
        if not self.Q:
            return np.zeros((self.nblk, 0), np.int32)
        return np.ascontiguousarray(np.concatenate(self.Q, axis=1))

    def dims(self) -> np.ndarray:
        lay = self._packed_layout()
        if lay is not None:
            return lay[2].copy()
        r = self.rank
        return np.array([d.shape[:r] for d in self.data], dtype=np.int64).reshape(len(self.data), r)

    def om(self) -> np.ndarray:
        """per-block outer multiplicity (size of the trailing data dimension
        beyond the rank, 1 if there is none)"""
        lay = self._packed_layout()
        if lay is not None:
            return lay[3].copy()
        r = self.rank
        return np.array([d.shape[r] if d.ndim > r else 1 for d in self.data], dtype=np.int32)

    def _set_packed(self, buf: np.ndarray, offs: np.ndarray, dims: np.ndarray, om: np.ndarray):
        """record that the blocks are the column-major views of one float64
        buffer, block i at element offset offs[i] (the form the library's
        results and with_packed produce): the marshalling then takes the
        block addresses from the buffer instead of visiting every block"""
        self._pk = (tuple(self.data), buf, offs, dims, om)

    def _packed_layout(self):
        """(buf, offs, dims, om) while the blocks are still the views
        _set_packed recorded (same objects, same order), else None"""
        pk = getattr(self, "_pk", None)
        if pk is None:
            return None
        views = pk[0]
        if len(views) != len(self.data) or any(a is not b for a, b in zip(views, self.data)):
            self._pk = None
            return None
        return pk[1:]

    @classmethod
    def _from_parts(cls, Q, data, qtype, itags, cgw) -> "QSpace":
        """a QSpace from parts the library produced (int32 label matrices,
        column-major float64 / complex128 blocks of full rank): no per-block
        conversion"""
        T = cls.__new__(cls)
        T.Q = Q
        T.data = data
        T.info = {"qtype": qtype, "itags": list(itags), "cgw": cgw}
        return T

    def numel(self) -> int:
        return int(sum(d.size for d in self.data))

    def norm(self) -> float:
        return float(np.sqrt(sum(float(np.sum(np.abs(d) ** 2)) for d in self.data)))

    def is_complex(self) -> bool:
        return any(d.dtype.kind == "c" for d in self.data)

    def packed(self) -> np.ndarray:
        """blocks back to back, column-major (the device layout)."""
        if not self.data:
            return np.zeros(0)
        return np.concatenate([d.ravel(order="F") for d in self.data])

    def with_packed(self, v: np.ndarray) -> "QSpace":
        out, o = [], 0
        for d in self.data:
            out.append(np.asarray(v[o: o + d.size]).reshape(d.shape, order="F"))
            o += d.size
        T = QSpace(self.Q, out, self.qtype, self.itags, self.info["cgw"])
        v = np.asarray(v)
        if (v.dtype == np.float64 and v.ndim == 1 and v.flags.c_contiguous
                and all(a.base is v or a.size == 0 for a in out) and all(a is b for a, b in zip(out, T.data))):
            offs = np.zeros(len(out) + 1, dtype=np.int64)
            np.cumsum([a.size for a in out], out=offs[1:])
            T._set_packed(v, offs, self.dims(), self.om())
        return T

    def copy(self) -> "QSpace":
        return QSpace([q.copy() for q in self.Q], [d.copy(order="F") for d in self.data], self.qtype,
                      list(self.itags), None if self.info["cgw"] is None else self.info["cgw"].copy())

    def __repr__(self):
        return f"QSpace(rank={self.rank}, nblk={self.nblk}, qtype='{self.qtype}', itags={self.itags})"


def _addr(a: np.ndarray) -> int:
    """data address of a Fortran-contiguous array: a ctypes view of its
    C-contiguous transpose (same first element) costs about 40% of an
    __array_interface__ lookup; read-only and empty arrays take the lookup"""
    try:
        return C.addressof(C.c_char.from_buffer(a.T))
    except (TypeError, ValueError):
        return a.__array_interface__["data"][0]


def to_c(T: QSpace, with_data: bool = True):
    """QSpace -> (QsxTensor, keepalive list)."""
    keep = []
    n, r = T.nblk, T.rank
    Q = np.ascontiguousarray(T.QIDX, dtype=np.int32).reshape(-1)
    if Q.size == 0:
        Q = np.zeros(1, np.int32)
    dims = np.ascontiguousarray(T.dims(), dtype=np.int64).reshape(-1)
    if dims.size == 0:
        dims = np.zeros(1, np.int64)
    keep += [Q, dims]
    ct = QsxTensor()
    ct.nblk, ct.rank, ct.qdim = n, r, T.qdim
    ct.Q = Q.ctypes.data_as(C.POINTER(C.c_int32))
    ct.dims = dims.ctypes.data_as(C.POINTER(C.c_int64))
    ct.data_im = None
    lay = T._packed_layout() if with_data else None
    if lay is not None and lay[0].dtype == np.float64:
        # the blocks are views of one buffer: addresses from the offsets
        buf, offs = lay[0], lay[1]
        addrs = (np.uint64(buf.__array_interface__["data"][0]) +
                 8 * offs[:-1].astype(np.uint64) if n else np.zeros(1, np.uint64)).astype(np.uintp)
        keep += [addrs, buf, T.data]
        ct.data = addrs.ctypes.data_as(C.POINTER(C.POINTER(C.c_double)))
    elif with_data:
        # block addresses gathered into one array (no per-block ctypes
        # objects); a block that is not column-major float64 gets a copy
        cplx = T.is_complex()
        addrs = np.zeros(max(1, n), dtype=np.uintp)
        aim = np.zeros(max(1, n), dtype=np.uintp) if cplx else None
        for i, d in enumerate(T.data):
            if cplx:
                a = np.asfortranarray(np.real(d), dtype=np.float64)
                b = np.asfortranarray(np.imag(d), dtype=np.float64)
                keep += [a, b]
                aim[i] = _addr(b)
            else:
                a = d if (d.dtype == np.float64 and d.flags.f_contiguous) else np.asfortranarray(d, dtype=np.float64)
                if a is not d:
                    keep.append(a)
            addrs[i] = _addr(a)
        keep += [addrs, T.data]
        ct.data = addrs.ctypes.data_as(C.POINTER(C.POINTER(C.c_double)))
        if cplx:
            keep.append(aim)
            ct.data_im = aim.ctypes.data_as(C.POINTER(C.POINTER(C.c_double)))
    else:
        ct.data = None
    if T.itags:
        it = np.array([pack_itag(t) for t in T.itags], dtype=np.uint64)
        keep.append(it)
        ct.itags = it.ctypes.data_as(C.POINTER(C.c_uint64))
    else:
        ct.itags = None
    qt = (T.qtype or "").encode()
    keep.append(qt)
    ct.qtype = qt
    if T.info.get("cgw") is not None:
        cg = np.ascontiguousarray(T.info["cgw"], dtype=np.float64).reshape(-1)
        keep.append(cg)
        ct.cgw = cg.ctypes.data_as(C.POINTER(C.c_double))
    else:
        ct.cgw = None
    if lay is not None:
        om = lay[3]
    else:
        om = T.om() if n and r and any(d.ndim > r for d in T.data) else np.ones(0, np.int32)
    if om.size and om.max() > 1:
        keep.append(om)
        ct.om = om.ctypes.data_as(C.POINTER(C.c_int32))
    else:
        ct.om = None
    return ct, keep


def from_out(o: QsxTensorOut, with_data: bool = True) -> QSpace:
    n, r, q = o.nblk, o.rank, o.qdim
    Qflat = np.ctypeslib.as_array(o.Q, shape=(max(1, n * r * q),))[: n * r * q].copy().reshape(n, r * q)
    dims = np.ctypeslib.as_array(o.dims, shape=(max(1, n * r),))[: n * r].copy().reshape(n, r)
    offs = np.ctypeslib.as_array(o.offs, shape=(n + 1,)).copy()
    data = []
    dl = dims.tolist()
    if o.om:
        om = np.ctypeslib.as_array(o.om, shape=(max(1, n),))[:n].tolist()
        dl = [d + [m] if m > 1 else d for d, m in zip(dl, om)]
    if with_data and o.data:
        # one copy of the whole result; the blocks are column-major views of it
        tot = int(offs[-1])
        buf = np.ctypeslib.as_array(o.data, shape=(max(1, tot),))[:tot].copy()
        if o.data_im:
            buf = buf + 1j * np.ctypeslib.as_array(o.data_im, shape=(max(1, tot),))[:tot]
        ol = offs.tolist()
        data = [buf[ol[i]: ol[i + 1]].reshape(dl[i] if r else (), order="F") for i in range(n)]
        packed = (buf, offs, dims) if buf.dtype == np.float64 else None
    else:
        packed = None
        for i in range(n):
            data.append(np.zeros(tuple(dl[i]), order="F"))
    itags = [unpack_itag(int(o.itags[i])) for i in range(r)] if o.itags else []
    if itags and all(t == "" for t in itags):
        itags = []
    cgw = None
    if o.nsym and o.cgw:
        cgw = np.ctypeslib.as_array(o.cgw, shape=(max(1, n * o.nsym),))[: n * o.nsym].copy().reshape(n, o.nsym)
    Qlist = [np.ascontiguousarray(Qflat[:, k * q:(k + 1) * q]) for k in range(r)]
    if not r:
        return _scalar_qspace(data, o.qtype.decode(), cgw)
    T = QSpace._from_parts(Qlist, data, o.qtype.decode(), itags, cgw)
    if packed is not None:
        om = (np.ctypeslib.as_array(o.om, shape=(max(1, n),))[:n].astype(np.int32) if o.om
              else np.ones(n, np.int32))
        T._set_packed(packed[0], packed[1], packed[2], om)
    return T


def _scalar_qspace(data, qtype, cgw):
    T = QSpace.__new__(QSpace)
    T.Q = []
    T.data = [np.asarray(d, dtype=np.float64) for d in data]
    T.info = {"qtype": qtype, "itags": [], "cgw": cgw}
    return T
