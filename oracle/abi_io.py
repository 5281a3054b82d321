"""The oracle's own encoder / decoder of the C-ABI tensor structs.

TEST INFRASTRUCTURE ONLY: used by oracle.py to hand QSpaces to the CPU
oracle (liboracle.so) and read its results, so that the oracle shares no
marshalling code with the product (qspace_amd/qspace.py): a layout or
itag-packing mistake there cannot hide by hitting both sides.  Written from
the struct definitions of include/qsx.h and the itag layout of the
reference (itag_, Source/wbindex.hh:632-700: up to 8 characters, little
endian in one 64-bit word, the conj flag -- an outgoing leg, written with a
trailing '*' -- in bit 7 of the first byte).  QSpace is only the container
(labels, blocks, info).
"""
from __future__ import annotations

import ctypes as C
import struct

import numpy as np

from qspace_amd.qspace import QSpace

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_f64p = C.POINTER(C.c_double)


class Tensor(C.Structure):
    """qsx_tensor"""
    _fields_ = [("nblk", C.c_int32), ("rank", C.c_int32), ("qdim", C.c_int32), ("Q", _i32p), ("dims", _i64p),
                ("data", C.POINTER(_f64p)), ("itags", C.POINTER(C.c_uint64)), ("qtype", C.c_char_p),
                ("cgw", _f64p), ("data_im", C.POINTER(_f64p)), ("om", _i32p)]


class TensorOut(C.Structure):
    """qsx_tensor_out"""
    _fields_ = [("nblk", C.c_int32), ("rank", C.c_int32), ("qdim", C.c_int32), ("nsym", C.c_int32),
                ("Q", _i32p), ("dims", _i64p), ("offs", _i64p), ("data", _f64p), ("itags", C.POINTER(C.c_uint64)),
                ("cgw", _f64p), ("qtype", C.c_char * 32), ("npairs", C.c_int64), ("useful_flops", C.c_double),
                ("gpu_ms", C.c_double), ("data_im", _f64p), ("om", _i32p)]


def itag_word(tag: str) -> int:
    """'name' / 'name*' -> the 64-bit itag word"""
    name = tag.rstrip("*")
    conj = (len(tag) - len(name)) % 2
    raw = name.encode("ascii")[:8] + b"\0" * 8
    (w,) = struct.unpack("<Q", raw[:8])
    return w | 0x80 if conj else w


def itag_str(w: int) -> str:
    raw = struct.pack("<Q", int(w) & ~0x80)
    name = raw.split(b"\0", 1)[0].decode("ascii", errors="replace")
    return name + ("*" if int(w) & 0x80 else "")


def encode(T: QSpace, with_data: bool = True):
    """QSpace -> (Tensor, arrays to keep alive)"""
    n, r = len(T.data), len(T.Q)
    q = T.Q[0].shape[1] if r else 1
    rec = np.zeros((max(n, 1), max(r * q, 1)), dtype=np.int32)
    for k in range(r):
        for c in range(q):
            rec[:n, k * q + c] = np.asarray(T.Q[k])[:, c]
    dims = np.ones((max(n, 1), max(r, 1)), dtype=np.int64)
    om = np.ones(max(n, 1), dtype=np.int32)
    for i, d in enumerate(T.data):
        shp = np.shape(d)
        for k in range(r):
            dims[i, k] = shp[k] if k < len(shp) else 1
        if len(shp) > r:
            om[i] = shp[r]
    keep = [rec, dims, om]
    t = Tensor()
    t.nblk, t.rank, t.qdim = n, r, q
    t.Q = rec.ctypes.data_as(_i32p)
    t.dims = dims.ctypes.data_as(_i64p)
    t.om = om.ctypes.data_as(_i32p) if n and om.max() > 1 else None
    t.data = t.data_im = None
    if with_data and n:
        cplx = any(np.iscomplexobj(d) for d in T.data)
        re_ptr = (_f64p * n)()
        im_ptr = (_f64p * n)() if cplx else None
        for i, d in enumerate(T.data):
            a = np.array(np.real(d), dtype=np.float64, order="F", copy=True)
            keep.append(a)
            re_ptr[i] = a.ctypes.data_as(_f64p)
            if cplx:
                b = np.array(np.imag(d), dtype=np.float64, order="F", copy=True)
                keep.append(b)
                im_ptr[i] = b.ctypes.data_as(_f64p)
        keep.append(re_ptr)
        t.data = C.cast(re_ptr, C.POINTER(_f64p))
        if cplx:
            keep.append(im_ptr)
            t.data_im = C.cast(im_ptr, C.POINTER(_f64p))
    tags = T.info.get("itags") or []
    if tags:
        w = (C.c_uint64 * r)(*[itag_word(x) for x in tags])
        keep.append(w)
        t.itags = C.cast(w, C.POINTER(C.c_uint64))
    else:
        t.itags = None
    qt = (T.info.get("qtype") or "").encode("ascii")
    keep.append(qt)
    t.qtype = qt
    cgw = T.info.get("cgw")
    if cgw is not None:
        cg = np.array(cgw, dtype=np.float64, order="C", copy=True).ravel()
        keep.append(cg)
        t.cgw = cg.ctypes.data_as(_f64p)
    else:
        t.cgw = None
    return t, keep


def decode(o: TensorOut, with_data: bool = True) -> QSpace:
    """a result struct -> QSpace (blocks copied out)"""
    n, r, q, ns = o.nblk, o.rank, o.qdim, o.nsym
    # records: n x r x q ints, block-major, then leg, then symmetry
    rec = np.ctypeslib.as_array(o.Q, shape=(max(n * r * q, 1),))[: n * r * q].reshape(n, r, q) if n * r else \
        np.zeros((n, r, q), np.int32)
    Q = [np.array(rec[:, k, :], dtype=np.int32) for k in range(r)]
    offs = np.ctypeslib.as_array(o.offs, shape=(n + 1,)).astype(np.int64)
    tot = int(offs[-1])
    buf = None
    if with_data and o.data:
        buf = np.array(np.ctypeslib.as_array(o.data, shape=(max(tot, 1),))[:tot])
        if o.data_im:
            buf = buf + 1j * np.ctypeslib.as_array(o.data_im, shape=(max(tot, 1),))[:tot]
    data = []
    for i in range(n):
        shp = [int(o.dims[i * r + k]) for k in range(r)]
        if o.om and o.om[i] > 1:
            shp.append(int(o.om[i]))
        v = buf[offs[i]: offs[i + 1]] if buf is not None else np.zeros(offs[i + 1] - offs[i])
        data.append(np.array(v).reshape(shp, order="F") if r else np.array(v).reshape(()))
    tags = [itag_str(o.itags[k]) for k in range(r)] if o.itags else []
    if tags and all(x == "" for x in tags):
        tags = []
    cgw = None
    if ns and o.cgw:
        cgw = np.array([[o.cgw[i * ns + j] for j in range(ns)] for i in range(n)], dtype=np.float64).reshape(n, ns)
    qtype = o.qtype.decode("ascii", errors="replace")
    if r:
        return QSpace(Q, data, qtype, tags, cgw)
    T = QSpace.__new__(QSpace)
    T.Q = []
    T.data = data
    T.info = {"qtype": qtype, "itags": [], "cgw": cgw}
    return T
