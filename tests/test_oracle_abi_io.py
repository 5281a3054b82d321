"""The oracle's own struct encoder / decoder (oracle/abi_io.py) against the
itag layout of the reference and against the product's marshalling (CPU
only): the same QSpace encoded by both gives byte-identical label records,
dims, itag words, weights, outer multiplicities and block contents; a result
struct decoded by both gives the same QSpace.  A layout mistake on either
side fails here instead of hitting both the product and its checker."""
import ctypes as C

import numpy as np
import pytest

import abi_io as IO
import oracle as O
import qspace_amd as Q
from qspace_amd import gen
from qspace_amd.qspace import QsxTensorOut, from_out, to_c


def test_itag_words_known_answers():
    # itag_ (Source/wbindex.hh:632-700): 8 chars little endian, conj = bit 7
    # of the first byte
    assert IO.itag_word("L") == 0x4C
    assert IO.itag_word("L*") == 0xCC
    assert IO.itag_word("abc") == 0x636261
    assert IO.itag_word("s1*") == (ord("s") | 0x80) | (ord("1") << 8)
    assert IO.itag_word("abcdefgh") == int.from_bytes(b"abcdefgh", "little")
    for t in ("L", "L*", "op12*", "x", ""):
        assert IO.itag_str(IO.itag_word(t)) == t
        assert IO.itag_word(t) == Q.pack_itag(t)


def _tensors():
    A = gen.u1_mps(seed=3, Lq=(-2, 0, 2), Rq=(-3, -1, 1, 3), dmin=1, dmax=3)
    S, L = gen.su2_mps(seed=4, tS=tuple(range(5)), dmin=1, dmax=3)
    W = gen.su2_om_tensor(5, [[(0, 2), (1, 1), (2, 1)]] * 4, ["a", "b", "c*", "d*"])
    Z = Q.QSpace(S.Q, [d * (1 + 0.5j) for d in S.data], S.qtype, S.itags)
    Wc = Q.QSpace(S.Q, S.data, S.qtype, S.itags, np.full((S.nblk, 1), 0.75))
    return [A, S, L, W, Z, Wc]


def _arr(p, n, dt):
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(dt)), shape=(n,)).copy() if p else None


@pytest.mark.parametrize("k", range(12))
def test_encoders_agree(k):
    """the product's and the oracle's encoders give the same struct; cases
    6.. are the same tensors re-made with with_packed (blocks as views of
    one buffer: the product takes their addresses from the offsets)"""
    ts = _tensors()
    T = ts[k % len(ts)]
    if k >= len(ts):
        if T.is_complex():
            pytest.skip("complex blocks are not views of one float64 buffer")
        T = T.with_packed(np.random.RandomState(k).standard_normal(T.numel()))
        assert T._packed_layout() is not None
    a, ka = to_c(T)
    b, kb = IO.encode(T)
    n, r, q = T.nblk, T.rank, T.qdim
    assert (a.nblk, a.rank, a.qdim) == (b.nblk, b.rank, b.qdim) == (n, r, q)
    assert np.array_equal(_arr(a.Q, n * r * q, C.c_int32), _arr(b.Q, n * r * q, C.c_int32))
    assert np.array_equal(_arr(a.dims, n * r, C.c_int64), _arr(b.dims, n * r, C.c_int64))
    assert np.array_equal(_arr(a.itags, r, C.c_uint64), _arr(b.itags, r, C.c_uint64))
    assert a.qtype == b.qtype
    ca, cb = _arr(a.cgw, n, C.c_double), _arr(b.cgw, n, C.c_double)
    assert (ca is None and cb is None) or np.array_equal(ca, cb)
    oa, ob = _arr(a.om, n, C.c_int32), _arr(b.om, n, C.c_int32)
    assert (oa is None and ob is None) or np.array_equal(oa, ob)
    for i, d in enumerate(T.data):
        m = d.size
        assert np.array_equal(_arr(a.data[i], m, C.c_double), _arr(b.data[i], m, C.c_double))
        if np.iscomplexobj(d):
            assert np.array_equal(_arr(a.data_im[i], m, C.c_double), _arr(b.data_im[i], m, C.c_double))
    # the labels the oracle encodes are the QSpace's, leg by leg
    rec = _arr(b.Q, n * r * q, C.c_int32).reshape(n, r, q)
    for kk in range(r):
        assert np.array_equal(rec[:, kk, :], T.Q[kk])


def test_decoders_agree():
    A = gen.u1_mps(seed=3, Lq=(-2, 0, 2), Rq=(-3, -1, 1, 3), dmin=1, dmax=3)
    L = O.lib()
    ca, ka = IO.encode(A)
    ia = np.array([0, 1], np.int32)
    out = C.POINTER(IO.TensorOut)()
    err = C.create_string_buffer(256)
    rc = L.oracle_contract(C.byref(ca), ia.ctypes.data_as(C.POINTER(C.c_int32)), 2, 1, C.byref(ca),
                           ia.ctypes.data_as(C.POINTER(C.c_int32)), 2, 0, None, 0, None, None, 1, C.byref(out), err,
                           256)
    assert rc == 0, err.value
    try:
        X = IO.decode(out.contents)
        Y = from_out(C.cast(out, C.POINTER(QsxTensorOut)).contents)
    finally:
        L.oracle_tensor_free(out)
    assert X.itags == Y.itags == ["R", "R*"]
    assert all(np.array_equal(x, y) for x, y in zip(X.Q, Y.Q))
    assert all(np.array_equal(x, y) for x, y in zip(X.data, Y.data))
    # a decoded result is one buffer: the product's encoder takes the block
    # addresses from its offsets, and they are the blocks' own
    assert Y._packed_layout() is not None
    a, ka = to_c(Y)
    for i, d in enumerate(Y.data):
        assert C.cast(a.data[i], C.c_void_p).value == d.__array_interface__["data"][0]


def test_packed_layout_follows_the_blocks():
    """the buffer layout recorded by with_packed is used only while the
    blocks are the recorded views: replacing a block drops it, and the
    struct then carries the new block's address"""
    T = gen.u1_mps(seed=5, Lq=(-2, 0, 2), Rq=(-3, -1, 1, 3), dmin=2, dmax=4)
    v = np.arange(T.numel(), dtype=np.float64)
    P = T.with_packed(v)
    lay = P._packed_layout()
    assert lay is not None and lay[0] is v
    assert np.array_equal(P.dims(), T.dims()) and np.array_equal(P.om(), T.om())
    a, ka = to_c(P)
    for i, d in enumerate(P.data):
        assert C.cast(a.data[i], C.c_void_p).value == d.__array_interface__["data"][0]
    new = np.full(P.data[1].shape, -1.0, order="F")
    P.data[1] = new
    assert P._packed_layout() is None
    b, kb = to_c(P)
    assert C.cast(b.data[1], C.c_void_p).value == new.__array_interface__["data"][0]
    assert np.array_equal(_arr(b.data[1], new.size, C.c_double), new.ravel(order="F"))
