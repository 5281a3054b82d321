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
  * reference's FP64 summation order / MKL dgemm rounding: parity unpinned
    (no reference build; tolerance 1e-12 relative instead of bit equality)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from qspace_amd.qspace import QSpace, QsxTensor, QsxTensorOut, from_out, to_c

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


def contract(A: QSpace, ia, B: QSpace, ib, perm=None, conjA=0, conjB=0, fermA=None, fermB=None,
             skip_zero=1) -> QSpace:
    """0-based ia/ib; returns the oracle's contractQS result."""
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
