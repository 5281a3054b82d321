"""qsx_block_zero_flags_noalloc (no device allocation: offsets and flags in
the library's fixed per-device array) against qsx_block_zero_flags, through
the C-ABI, on two calls in a row -- the second with fewer blocks, so stale
flags of the first call would show -- and its size limit."""
import ctypes as C

import numpy as np
import pytest

import qspace_amd as Q

pytestmark = pytest.mark.gpu

MAXB = 65536


def _flags(fn, offs, dbuf, n):
    out = np.full(max(1, n), -7, np.int32)
    err = C.create_string_buffer(256)
    P = C.POINTER
    rc = fn(offs.ctypes.data_as(P(C.c_int64)), n, dbuf.ptr, out.ctypes.data_as(P(C.c_int32)), err, 256)
    Q._capi.check(rc, err)
    return out[:n]


def _case(rng, nblk, zero_frac, maxsize=3000):
    sizes = rng.randint(0, maxsize, nblk)
    sizes[rng.rand(nblk) < 0.05] = 0  # empty blocks
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    data = rng.standard_normal(int(offs[-1]) + 1)
    for i in np.where(rng.rand(nblk) < zero_frac)[0]:
        data[offs[i]:offs[i + 1]] = 0.0
        if offs[i + 1] > offs[i] and rng.rand() < 0.5:  # one non-zero at a random place
            data[offs[i] + rng.randint(offs[i + 1] - offs[i])] = 1e-300
    return offs, data


def test_noalloc_zero_flags_match(gpu):
    L = Q._capi.lib()
    rng = np.random.RandomState(12)
    for nblk in (5000, 700):  # second call smaller than the first
        offs, data = _case(rng, nblk, 0.4)
        d = Q.DeviceBuffer.from_array(data)
        ref = _flags(L.qsx_block_zero_flags, offs, d, nblk)
        got = _flags(L.qsx_block_zero_flags_noalloc, offs, d, nblk)
        print(f"noalloc zero flags: {nblk} blocks, {int(ref.sum())} zero", flush=True)
        assert np.array_equal(got, ref)
        want = np.array([not np.any(data[offs[i]:offs[i + 1]]) for i in range(nblk)], np.int32)
        assert np.array_equal(got, want)


def test_noalloc_zero_flags_limit(gpu):
    L = Q._capi.lib()
    rng = np.random.RandomState(3)
    offs, data = _case(rng, MAXB, 0.3, maxsize=64)  # the largest list it takes
    d = Q.DeviceBuffer.from_array(data)
    assert np.array_equal(_flags(L.qsx_block_zero_flags_noalloc, offs, d, MAXB),
                          _flags(L.qsx_block_zero_flags, offs, d, MAXB))
    offs2 = np.arange(MAXB + 2, dtype=np.int64)
    out = np.zeros(MAXB + 1, np.int32)
    err = C.create_string_buffer(256)
    rc = L.qsx_block_zero_flags_noalloc(offs2.ctypes.data_as(C.POINTER(C.c_int64)), MAXB + 1, d.ptr,
                                        out.ctypes.data_as(C.POINTER(C.c_int32)), err, 256)
    assert rc == 1 and b"maximum" in err.value
