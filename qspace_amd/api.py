"""Python mirror of the reference's contraction interface.

``contractQS(A, ica, B, icb [, perm] [, opts])`` follows the MATLAB usage #1
of bin/contractQS.m:20-191 (1-based numeric indices or strings like '13*';
trailing '*' = complex conjugate of that operand), and the cell usage #2
``contractQS([A, B], C, ...)`` contracts by itag auto-matching
(contractQS_itags, Source/contractQS.cc:634-808).  Computation always runs
through libqsx.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from .qspace import QSpace, QsxCell, QsxOpts, QsxTensor, QsxTensorOut, from_out, parse_ctridx, to_c

__all__ = ["contractQS", "permuteQS", "plusQS", "orthoQS", "ortho_args", "itag_match", "DeviceBuffer", "CtrPlan", "HPsi", "su2_cg", "su2_x3", "su2_om_basis",
           "su2_xom", "su2_canon_order", "make_opts"]


def make_opts(mode="batched", nslices=0, skip_zero=1, hconj=0, fermA=None, fermB=None, verbose=1):
    o = QsxOpts()
    o.mode = _capi.MODES[mode] if isinstance(mode, str) else int(mode)
    o.verbose = verbose
    o.nslices = nslices
    o.skip_zero = skip_zero
    o.hconj = hconj
    keep = []
    if fermA is not None:
        fa = np.ascontiguousarray(fermA, dtype=np.float64)
        keep.append(fa)
        o.fermA = fa.ctypes.data_as(C.POINTER(C.c_double))
    if fermB is not None:
        fb = np.ascontiguousarray(fermB, dtype=np.float64)
        keep.append(fb)
        o.fermB = fb.ctypes.data_as(C.POINTER(C.c_double))
    return o, keep


POLICY_KEYS = ("engine", "split_steps", "tile_overhead", "slots", "wave_max", "group_mfma", "split_mfma", "xcd_bins",
               "group_waves", "r9", "r10", "r11")
NPOL = len(POLICY_KEYS)
ENGINES = {"grouped": 0, "wave": 1}
_POLICY_DEFAULT = {"engine": 1, "split_steps": 16, "tile_overhead": 1024, "slots": 1024, "wave_max": 2,
                   "group_mfma": 64, "split_mfma": 1 << 30, "xcd_bins": 8, "group_waves": 4, "r9": 0,
                   "r10": 0, "r11": 0}


def _policy_ints(p: dict):
    """int32[8] exchange form of a work-decomposition policy (qsx.h)."""
    d = dict(_POLICY_DEFAULT)
    for k, v in p.items():
        if k not in d:
            raise KeyError(f"unknown policy key {k}")
        d[k] = ENGINES[v] if k == "engine" and isinstance(v, str) else int(v)
    return [d[k] for k in POLICY_KEYS]


def _policy_dict(v):
    """named form of an int32 policy vector (the reserved slots dropped)"""
    return {k: int(x) for k, x in zip(POLICY_KEYS[:9], v)}


def _ivec(v):
    a = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape(-1))
    return a if a.size else np.zeros(1, np.int32)


def _contract_pair(A, ia, conjA, B, ib, conjB, perm=None, mode="batched", **kw) -> QSpace:
    L = _capi.lib()
    ca, ka = to_c(A)
    cb, kb = to_c(B)
    opts, keep = make_opts(mode=mode, **kw)
    ia_ = _ivec(ia)
    ib_ = _ivec(ib)
    p = _ivec(perm if perm is not None else [])
    out = C.POINTER(QsxTensorOut)()
    err = C.create_string_buffer(1024)
    rc = L.qsx_contract(C.byref(ca), ia_.ctypes.data_as(C.POINTER(C.c_int32)), len(ia), int(conjA), C.byref(cb),
                        ib_.ctypes.data_as(C.POINTER(C.c_int32)), len(ib), int(conjB),
                        p.ctypes.data_as(C.POINTER(C.c_int32)), 0 if perm is None else len(perm), C.byref(opts),
                        C.byref(out), err, 1024)
    _capi.check(rc, err)
    try:
        T = from_out(out.contents)
        T.npairs = int(out.contents.npairs)
        T.flops = float(out.contents.useful_flops)
        T.gpu_ms = float(out.contents.gpu_ms)
    finally:
        L.qsx_tensor_free(out)
    return T


def contractQS(*args, mode="batched", **kw) -> QSpace:
    """contractQS(A, ica, B, icb [,perm]) or the cell form
    contractQS({A, opts.., B, opts..}, C, opts.. [,perm] [,'-v'|'-q'])."""
    # usage #1 when arguments 2 and 4 are contraction indices (isCtrIdx,
    # Source/contractQS.cc:369-373)
    if len(args) >= 4 and isinstance(args[0], QSpace) and isinstance(args[2], QSpace) and \
            _is_ctridx(args[1]) and _is_ctridx(args[3]):
        A, ica, B, icb = args[:4]
        perm = None
        if len(args) > 4 and args[4] is not None and not isinstance(args[4], str):
            perm = [int(x) - 1 for x in np.atleast_1d(args[4])]
        ia, ca = parse_ctridx(ica)
        ib, cb = parse_ctridx(icb)
        return _contract_pair(A, ia, ca, B, ib, cb, perm, mode=mode, **kw)
    # usage #2 (mexFunction, Source/contractQS.cc:364-433): trailing -v/-q,
    # then a trailing numeric vector that is a valid permutation is the
    # output permutation (a single index only at argument position < 2; a
    # later one is the preceding operand's option)
    args = list(args)
    while args and isinstance(args[-1], str) and args[-1] in ("-v", "-q"):
        args.pop()
    perm = None
    if len(args) >= 2 and _is_numvec(args[-1]):
        v = [int(x) for x in np.atleast_1d(args[-1])]
        if sorted(v) == list(range(1, len(v) + 1)) and (len(v) != 1 or len(args) - 1 < 2):
            perm = [x - 1 for x in v]
            args.pop()
    # the fold (itag matching, options, device-resident operands) is the
    # library's qsx_contract_cell
    L = _capi.lib()
    keep = []
    items, n = _cell_structs(args, {}, keep)
    opts, okeep = make_opts(mode=mode, **kw)
    p = _ivec(perm if perm is not None else [])
    out = C.POINTER(QsxTensorOut)()
    err = C.create_string_buffer(1024)
    rc = L.qsx_contract_cell(items, n, p.ctypes.data_as(C.POINTER(C.c_int32)), 0 if perm is None else len(perm),
                             C.byref(opts), C.byref(out), err, 1024)
    _capi.check(rc, err)
    return _out_to_qspace(out)


def _cell_structs(args, tensors, keep):
    """the argument list of the cell form as qsx_cell entries: a QSpace or a
    nested cell starts an entry, the strings and numeric vectors after it are
    its options; one QsxTensor per distinct QSpace (the library uploads each
    leaf once)"""
    entries = []
    for a in args:
        if isinstance(a, QSpace):
            if id(a) not in tensors:
                ct, k = to_c(a)
                tensors[id(a)] = C.pointer(ct)
                keep += [k, ct, a]
            entries.append({"T": tensors[id(a)], "sub": None, "opt": [], "idx": None})
        elif _is_cell(a):
            sub, n = _cell_structs(list(a), tensors, keep)
            entries.append({"T": None, "sub": (sub, n), "opt": [], "idx": None})
        else:
            if not entries:
                raise _capi.QsxError(1, "invalid QSpace cell-structure (option before any QSpace)")
            if isinstance(a, str):
                entries[-1]["opt"].append(a)
            elif _is_numvec(a):
                if entries[-1]["idx"] is not None:
                    raise _capi.QsxError(1, "cell-contraction option: ctr-idx already set")
                entries[-1]["idx"] = [int(x) for x in np.atleast_1d(a)]
            else:
                raise _capi.QsxError(1, f"unexpected cell-contraction input {type(a).__name__}")
    arr = (QsxCell * max(1, len(entries)))()
    for e, c in zip(entries, arr):
        if e["T"] is not None:
            c.T = e["T"]
        else:
            c.sub, c.nsub = e["sub"]
        _cell_options(c, e["opt"], e["idx"], keep)
    keep.append(arr)
    return arr, len(entries)


def _cell_options(c, opt, idx, keep):
    if opt:
        b = [o.encode() for o in opt]
        a = (C.c_char_p * len(b))(*b)
        keep += [b, a]
        c.opt, c.nopt = a, len(b)
    if idx is not None:
        v = _ivec(idx)
        keep.append(v)
        c.idx, c.nidx = v.ctypes.data_as(C.POINTER(C.c_int32)), len(idx)


def itag_match(A: QSpace, optA, B: QSpace, optB):
    """one step of the cell form's fold, structure only (qsx_itag_match, no
    device): the options of A and B (lists of option strings / one numeric
    index vector, as in the cell form) and the itag matching give (ia, ib,
    conjA, conjB), ia / ib 0-based"""
    L = _capi.lib()
    keep = []
    ca, k1 = to_c(A, with_data=False)
    cb, k2 = to_c(B, with_data=False)
    cells = (QsxCell * 2)()
    for c, o in zip(cells, (optA, optB)):
        strs = [x for x in o if isinstance(x, str)]
        nums = [x for x in o if not isinstance(x, str)]
        if len(nums) > 1:
            raise _capi.QsxError(1, "cell-contraction option: ctr-idx already set")
        _cell_options(c, strs, [int(x) for x in np.atleast_1d(nums[0])] if nums else None, keep)
    n = max(A.rank, B.rank, 1)
    ia, ib = np.zeros(n, np.int32), np.zeros(n, np.int32)
    nc, cA, cB = C.c_int32(), C.c_int32(), C.c_int32()
    err = C.create_string_buffer(512)
    P = C.POINTER
    rc = L.qsx_itag_match(C.byref(ca), cells[0].nopt, cells[0].opt, cells[0].idx, cells[0].nidx, C.byref(cb),
                          cells[1].nopt, cells[1].opt, cells[1].idx, cells[1].nidx,
                          ia.ctypes.data_as(P(C.c_int32)), ib.ctypes.data_as(P(C.c_int32)), C.byref(nc),
                          C.byref(cA), C.byref(cB), err, 512)
    _capi.check(rc, err)
    return list(ia[: nc.value]), list(ib[: nc.value]), cA.value, cB.value


def _is_ctridx(a) -> bool:
    if isinstance(a, QSpace) or _is_cell(a):
        return False
    try:
        idx, _ = parse_ctridx(a)
    except Exception:
        return False
    return len(idx) > 0


def _is_cell(a) -> bool:
    return isinstance(a, (list, tuple)) and len(a) > 0 and isinstance(a[0], (QSpace, list, tuple))


def _is_numvec(a) -> bool:
    if isinstance(a, (str, QSpace)) or _is_cell(a):
        return False
    try:
        v = np.atleast_1d(np.asarray(a))
    except (TypeError, ValueError):
        return False
    return v.ndim == 1 and v.size > 0 and np.issubdtype(v.dtype, np.number)


def _out_to_qspace(out) -> QSpace:
    try:
        T = from_out(out.contents)
        T.gpu_ms = float(out.contents.gpu_ms)
    finally:
        _capi.lib().qsx_tensor_free(out)
    return T


def permuteQS(A: QSpace, perm, conj: bool = False) -> QSpace:
    """permuteQS(A, perm [,'conj']) (Source/permuteQS.cc:25-87) on the device.

    ``perm`` is 0-based and may be shorter than the rank (completed with the
    identity); ``conj`` flips the conj flag of every leg."""
    L = _capi.lib()
    ca, ka = to_c(A)
    p = _ivec(list(perm))
    out = C.POINTER(QsxTensorOut)()
    err = C.create_string_buffer(1024)
    rc = L.qsx_permute(C.byref(ca), p.ctypes.data_as(C.POINTER(C.c_int32)), len(list(perm)), int(bool(conj)),
                       C.byref(out), err, 1024)
    _capi.check(rc, err)
    return _out_to_qspace(out)


def plusQS(A: QSpace, B: QSpace, bfac: float = 1.0, skip_zero: int = 1) -> QSpace:
    """plusQS(A, B [,bfac]) (Source/plusQS.cc:17-87): A + bfac*B on the device."""
    L = _capi.lib()
    ca, ka = to_c(A)
    cb, kb = to_c(B)
    opts, keep = make_opts(skip_zero=skip_zero)
    out = C.POINTER(QsxTensorOut)()
    err = C.create_string_buffer(1024)
    rc = L.qsx_plus(C.byref(ca), C.byref(cb), float(bfac), C.byref(opts), C.byref(out), err, 1024)
    _capi.check(rc, err)
    return _out_to_qspace(out)


STOL_SVD = 1e-8  # WB_STOL_SVD, Source/mpsortho.hh:4


def ortho_args(PSI: QSpace, K: int, Nkeep=None, stol=None):
    """(Nkeep, stol) as the orthoQS front end resolves them
    (MPS_ORTHO_1, Source/orthoQS.cc:137-170): stol defaults to
    WB_STOL_SVD = 1e-8, or to 0 when Nkeep > 0 is given without stol; a
    missing or negative Nkeep becomes get_Nkeep_estimate (orthoQS.cc:39-56):
    with d1 = prod of the leg dims of legs 0..K-1 and d2 that of legs
    K..r-1 (QSpace::getDim: per leg the sum of the block dims over its
    distinct labels, multiplets for non-abelian symmetries), min(d1, d2) for
    abelian symmetries and max(d1, d2) otherwise."""
    if stol is None:
        stol = 0.0 if (Nkeep is not None and Nkeep > 0) else STOL_SVD
    if Nkeep is None or Nkeep < 0:
        dims = PSI.dims()
        leg = []
        for k in range(PSI.rank):
            seen = {}
            for i in range(PSI.nblk):
                seen.setdefault(tuple(np.asarray(PSI.Q[k][i]).ravel()), int(dims[i, k]))
            leg.append(sum(seen.values()))
        d1 = int(np.prod(leg[:K], dtype=np.int64))
        d2 = int(np.prod(leg[K:], dtype=np.int64))
        abelian = all(x.strip() in ("", "A") for x in (PSI.qtype or "").split(","))
        Nkeep = min(d1, d2) if abelian else max(d1, d2)
    return int(Nkeep), float(stol)


def orthoQS(PSI: QSpace, K: int, Nkeep=None, stol=None, Nkmin: int = 0, itag: str = ""):
    """orthoQS(PSI, K, Nkeep, stol, Nkmin, itag) -- the truncation core of
    Source/mpsortho.cc:1031-1078 (blockSVD + dmrgTruncate) on the device,
    with the front end's defaults for Nkeep and stol (``ortho_args``).

    Returns (A1, A2, info): A1 = U with legs (PSI legs 0..K-1, new leg),
    A2 = V S sfac with legs (PSI legs K..r-1, new leg), PSI ~ A1 . A2 over
    the new leg; info holds svd (all singular values, descending), Nkeep,
    Ntot, svd2tr, sv2tot, sfac, flag, and the device time.  Real data;
    abelian symmetries, and SU(2) for K = 1 or r-1 splits of rank <= 3 and
    K = 2 of a rank-4 tensor (outer multiplicity) (info then holds
    svd_qdim, the multiplet dimension of each singular value)."""
    Nkeep, stol = ortho_args(PSI, K, Nkeep, stol)
    L = _capi.lib()
    ca, ka = to_c(PSI)
    o1 = C.POINTER(QsxTensorOut)()
    o2 = C.POINTER(QsxTensorOut)()
    oi = C.POINTER(_capi.QsxOrthoInfo)()
    err = C.create_string_buffer(1024)
    rc = L.qsx_ortho(C.byref(ca), int(K), int(Nkmin), int(Nkeep), float(stol), (itag or "").encode(),
                     C.byref(o1), C.byref(o2), C.byref(oi), err, 1024)
    _capi.check(rc, err)
    try:
        x = oi.contents
        info = dict(svd=np.ctypeslib.as_array(x.svd, shape=(max(1, x.ntot),))[:x.ntot].copy(), Nkeep=int(x.nkeep),
                    Ntot=int(x.ntot), Nkmin=int(x.nkmin), flag=int(x.flag), stol=float(x.stol),
                    svd2tr=float(x.svd2tr), sv2tot=float(x.sv2tot), sfac=float(x.sfac), gpu_ms=float(x.gpu_ms),
                    sweeps=int(x.sweeps), ngroups=int(x.ngroups))
        if any(s.strip() == "SU2" for s in (PSI.qtype or "").split(",")):
            info["svd_qdim"] = np.ctypeslib.as_array(x.svd_qdim, shape=(max(1, x.ntot),))[:x.ntot].copy()
    finally:
        L.qsx_ortho_info_free(oi)
    return _out_to_qspace(o1), _out_to_qspace(o2), info


class DeviceBuffer:
    """A device allocation of n doubles owned by libqsx."""

    def __init__(self, n: int):
        L = _capi.lib()
        self.n = int(n)
        self.ptr = C.c_void_p()
        self._release = L.qsx_free_device  # held so teardown at exit need not look up the module
        err = C.create_string_buffer(256)
        _capi.check(L.qsx_malloc_device(C.byref(self.ptr), 8 * max(1, self.n), err, 256), err)

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        if np.iscomplexobj(a):
            raise TypeError("DeviceBuffer holds real doubles (got complex data)")
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = cls(a.size)
        b.upload(a)
        return b

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=np.float64)
        _capi.check(_capi.lib().qsx_memcpy_h2d(self.ptr, a.ctypes.data, 8 * a.size, None), None)

    def download(self) -> np.ndarray:
        out = np.empty(self.n)
        _capi.check(_capi.lib().qsx_memcpy_d2h(out.ctypes.data, self.ptr, 8 * self.n, None), None)
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            self._release(self.ptr)
            self.ptr = None


class CtrPlan:
    """Device-resident contraction plan (structure once, execute many)."""

    def __init__(self, A: QSpace, ica, B: QSpace, icb, perm=None, mode="batched", **kw):
        L = _capi.lib()
        ia, ca = parse_ctridx(ica)
        ib, cb = parse_ctridx(icb)
        cA, self._ka = to_c(A, with_data=False)
        cB, self._kb = to_c(B, with_data=False)
        self._opts, self._okeep = make_opts(mode=mode, **kw)
        ia_ = _ivec(ia)
        ib_ = _ivec(ib)
        p = _ivec(perm if perm is not None else [])
        self.h = C.c_void_p()
        self._release = L.qsx_ctr_plan_destroy
        err = C.create_string_buffer(1024)
        rc = L.qsx_ctr_plan_create(C.byref(cA), ia_.ctypes.data_as(C.POINTER(C.c_int32)), len(ia), ca, C.byref(cB),
                                   ib_.ctypes.data_as(C.POINTER(C.c_int32)), len(ib), cb,
                                   p.ctypes.data_as(C.POINTER(C.c_int32)), 0 if perm is None else len(perm),
                                   C.byref(self._opts), C.byref(self.h), err, 1024)
        _capi.check(rc, err)
        out = C.POINTER(QsxTensorOut)()
        L.qsx_ctr_plan_output(self.h, C.byref(out))
        self.out = from_out(out.contents, with_data=False)
        L.qsx_tensor_free(out)
        self.flops = L.qsx_ctr_plan_flops(self.h)
        self.nout = self.out.numel()
        self._rank_out = A.rank + B.rank - 2 * len(ia)
        self._qdim = A.qdim

    def pairs(self):
        L = _capi.lib()
        n = C.c_int64()
        m = C.c_int64()
        L.qsx_ctr_plan_pairs(self.h, C.byref(n), C.byref(m), None, None, None, None)
        ia = np.zeros(max(1, n.value), np.int32)
        ib = np.zeros(max(1, n.value), np.int32)
        af = np.zeros(max(1, n.value))
        seg = np.zeros(m.value + 1, np.int64)
        P = C.POINTER
        L.qsx_ctr_plan_pairs(self.h, C.byref(n), C.byref(m), ia.ctypes.data_as(P(C.c_int32)),
                             ib.ctypes.data_as(P(C.c_int32)), af.ctypes.data_as(P(C.c_double)),
                             seg.ctypes.data_as(P(C.c_int64)))
        return ia[: n.value], ib[: n.value], af[: n.value], seg

    def idxset(self):
        """contract_getIdxSet's result as the plan holds it: (Ia, Ib, afac,
        seg, QC) over all label-matched pairs, afac before NormCGW"""
        L = _capi.lib()
        n = C.c_int64()
        m = C.c_int64()
        _capi.check(L.qsx_ctr_plan_idxset(self.h, C.byref(n), C.byref(m), None, None, None, None, None), None)
        rc = self._rank_out
        q = self._qdim
        ia = np.zeros(max(1, n.value), np.int32)
        ib = np.zeros(max(1, n.value), np.int32)
        af = np.zeros(max(1, n.value))
        seg = np.zeros(m.value + 1, np.int64)
        qc = np.zeros(max(1, m.value * rc * q), np.int32)
        P = C.POINTER
        _capi.check(L.qsx_ctr_plan_idxset(self.h, C.byref(n), C.byref(m), ia.ctypes.data_as(P(C.c_int32)),
                                          ib.ctypes.data_as(P(C.c_int32)), af.ctypes.data_as(P(C.c_double)),
                                          seg.ctypes.data_as(P(C.c_int64)), qc.ctypes.data_as(P(C.c_int32))), None)
        return ia[: n.value], ib[: n.value], af[: n.value], seg, qc[: m.value * rc * q].reshape(m.value, rc * q)

    def execute(self, dA: DeviceBuffer, dB: DeviceBuffer, dC: DeviceBuffer):
        _capi.check(_capi.lib().qsx_ctr_execute(self.h, dA.ptr, dB.ptr, dC.ptr, None), None)

    def set_policy(self, **policy):
        """work decomposition of the batched GEMM (keys of POLICY_KEYS)"""
        v = (C.c_int32 * NPOL)(*_policy_ints(policy))
        _capi.check(_capi.lib().qsx_ctr_plan_set_policy(self.h, v), None)

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self._release(self.h)
            self.h = None


class HPsi:
    """Two-site H*psi = contractQS({HL,psi},HR) [+ hconj term], device resident."""

    def __init__(self, HL: QSpace, psi: QSpace, HR: QSpace, mode="batched", hconj=1, fermA=None, nslices=0,
                 verbose=0):
        L = _capi.lib()
        cL, self._k1 = to_c(HL)
        cP, self._k2 = to_c(psi, with_data=False)
        cR, self._k3 = to_c(HR)
        self._opts, self._okeep = make_opts(mode=mode, hconj=hconj, fermA=fermA, nslices=nslices, skip_zero=0,
                                            verbose=verbose)
        self.h = C.c_void_p()
        self._release = L.qsx_hpsi_plan_destroy
        err = C.create_string_buffer(1024)
        rc = L.qsx_hpsi_plan_create(C.byref(cL), C.byref(cP), C.byref(cR), C.byref(self._opts), C.byref(self.h),
                                    err, 1024)
        _capi.check(rc, err)
        self.psi = psi
        self.N = int(L.qsx_hpsi_size(self.h))
        self.flops = float(L.qsx_hpsi_flops(self.h))
        self.bytes = float(L.qsx_hpsi_bytes(self.h))

    @classmethod
    def pseudo(cls, HKL, psi: QSpace, HKR, OPL, OPR, hconj=None, mode="batched", verbose=0) -> "HPsi":
        """H*psi of the pseudo MPO (get_HPsi.m:62-83): HK_L.psi + psi.HK_R +
        sum_j contractQS({OP_L[j],'*',psi},OP_R[j]) [+ contractQS({OP_L[j],psi},
        OP_R[j],'*') where hconj[j]]; HKL / HKR may be None, hconj defaults to
        1 for every operator."""
        L = _capi.lib()
        if len(OPL) != len(OPR):
            raise ValueError("OP_L and OP_R lists differ in length")
        self = cls.__new__(cls)
        keep = []
        cL = cR = None
        if HKL is not None:
            cL, k = to_c(HKL)
            keep.append(k)
        if HKR is not None:
            cR, k = to_c(HKR)
            keep.append(k)
        n = len(OPL)
        arrL = (QsxTensor * max(1, n))()
        arrR = (QsxTensor * max(1, n))()
        for j in range(n):
            arrL[j], k1 = to_c(OPL[j])
            arrR[j], k2 = to_c(OPR[j])
            keep += [k1, k2]
        hc = np.ascontiguousarray([1] * n if hconj is None else [int(x) for x in hconj], dtype=np.int32)
        if hc.size != n:
            raise ValueError("hconj needs one flag per operator")
        cP, kp = to_c(psi, with_data=False)
        keep.append(kp)
        self._opts, self._okeep = make_opts(mode=mode, skip_zero=0, verbose=verbose)
        self._k1 = keep
        self.h = C.c_void_p()
        self._release = L.qsx_hpsi_plan_destroy
        err = C.create_string_buffer(1024)
        P = C.POINTER
        rc = L.qsx_hpsi_plan_create_pseudo(C.byref(cL) if cL is not None else None,
                                           C.byref(cR) if cR is not None else None, n, arrL, arrR,
                                           hc.ctypes.data_as(P(C.c_int32)), C.byref(cP), C.byref(self._opts),
                                           C.byref(self.h), err, 1024)
        _capi.check(rc, err)
        self.psi = psi
        self.N = int(L.qsx_hpsi_size(self.h))
        self.flops = float(L.qsx_hpsi_flops(self.h))
        self.bytes = float(L.qsx_hpsi_bytes(self.h))
        return self

    def stage_info(self) -> dict:
        """per stage: useful flops, algorithmic bytes, work items, pairs, the
        GEMM kernel that runs it, its workgroups and reduction workgroups"""
        v = (C.c_double * 20)()
        _capi.check(_capi.lib().qsx_hpsi_stage_info(self.h, v), None)
        kern = []
        for t in (0, 1):
            if int(v[8 + t]) == 0:
                kern.append("k_grouped")
            elif int(v[8 + t]) == 2:
                kern.append("k_split_gemm")
            else:
                kern.append("k_wave_s" if int(v[12 + t]) <= 2 else "k_wave_l")
        return {"flops": (v[0], v[1]), "bytes": (v[2], v[3]), "tiles": (int(v[4]), int(v[5])),
                "pairs": (int(v[6]), int(v[7])), "kernel": tuple(kern), "grid": (int(v[10]), int(v[11])),
                "reduce_grid": (int(v[14]), int(v[15])), "split_exec_flops": (v[16], v[17]),
                "mfma_exec_flops": (v[18], v[19])}

    def autotune(self, reps: int = 5):
        """time candidate work decompositions of both stages and keep the
        fastest: {"stage_ms": (t1, t2), "policy": (stage-1 dict, stage-2 dict)}"""
        ms = (C.c_double * 2)()
        pol = (C.c_int32 * (2 * NPOL))()
        _capi.check(_capi.lib().qsx_hpsi_autotune(self.h, reps, ms, pol), None)
        return {"stage_ms": (ms[0], ms[1]), "policy": (_policy_dict(pol[:NPOL]), _policy_dict(pol[NPOL:]))}

    def set_policy(self, stage1: dict, stage2: dict):
        """work decomposition per stage (keys of POLICY_KEYS)"""
        v = (C.c_int32 * (2 * NPOL))(*(_policy_ints(stage1) + _policy_ints(stage2)))
        _capi.check(_capi.lib().qsx_hpsi_set_policy(self.h, v), None)

    def weights(self) -> np.ndarray:
        w = np.zeros(self.N)
        _capi.check(_capi.lib().qsx_hpsi_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_double))), None)
        return w

    def use_graph(self, on: bool):
        _capi.lib().qsx_hpsi_use_graph(self.h, int(on))

    def apply_dev(self, dpsi: DeviceBuffer, dhpsi: DeviceBuffer):
        _capi.check(_capi.lib().qsx_hpsi_apply(self.h, dpsi.ptr, dhpsi.ptr, None), None)

    def apply(self, v: np.ndarray) -> np.ndarray:
        d = DeviceBuffer.from_array(v)
        o = DeviceBuffer(self.N)
        self.apply_dev(d, o)
        return o.download()

    def last_timing(self):
        """(stage 1, stage 2) grouped-GEMM ms of the last direct apply"""
        ms = (C.c_double * 2)()
        _capi.check(_capi.lib().qsx_hpsi_last_timing(self.h, ms), None)
        return tuple(ms)

    def stage_ms(self, stage: int, reps: int, dpsi: DeviceBuffer, dhpsi: DeviceBuffer) -> float:
        """mean device ms per launch of one stage over reps back-to-back
        launches (two HIP events on the launch stream)"""
        ms = C.c_double()
        _capi.check(_capi.lib().qsx_hpsi_stage_ms(self.h, int(stage), int(reps), dpsi.ptr, dhpsi.ptr, C.byref(ms)),
                    None)
        return ms.value

    def lanczos(self, psi0: np.ndarray, niter: int):
        return self._eig(_capi.lib().qsx_lanczos, psi0, niter)

    def davidson(self, psi0: np.ndarray, niter: int):
        return self._eig(_capi.lib().qsx_davidson, psi0, niter)

    def _eig(self, fn, psi0, niter):
        p0 = np.ascontiguousarray(psi0, dtype=np.float64)
        e0 = C.c_double()
        ritz = np.zeros(niter)
        gs = np.zeros(self.N)
        err = C.create_string_buffer(512)
        P = C.POINTER
        rc = fn(self.h, p0.ctypes.data_as(P(C.c_double)), niter, C.byref(e0), ritz.ctypes.data_as(P(C.c_double)),
                gs.ctypes.data_as(P(C.c_double)), err, 512)
        _capi.check(rc, err)
        return e0.value, ritz, gs

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            self._release(self.h)
            self.h = None


def su2_cg(tj1, tm1, tj2, tm2, tj3, tm3) -> float:
    """<j1 m1; j2 m2 | j3 m3> (two-j / two-m ints), exact-arithmetic host engine."""
    return float(_capi.lib().qsx_su2_cg(tj1, tm1, tj2, tm2, tj3, tm3))


def su2_x3(tja, dira, tjb, dirb, ia, ib) -> float:
    L = _capi.lib()
    P = C.POINTER
    a, da, b, db, ca, cb = (_ivec(v) for v in (tja, dira, tjb, dirb, ia, ib))
    out = C.c_double()
    err = C.create_string_buffer(256)
    rc = L.qsx_su2_x3(len(tja), a.ctypes.data_as(P(C.c_int32)), da.ctypes.data_as(P(C.c_int32)), len(tjb),
                      b.ctypes.data_as(P(C.c_int32)), db.ctypes.data_as(P(C.c_int32)), ca.ctypes.data_as(P(C.c_int32)),
                      cb.ctypes.data_as(P(C.c_int32)), len(ia), C.byref(out), err, 256)
    _capi.check(rc, err)
    return out.value


def su2_om_basis(tj, dirs) -> np.ndarray:
    """outer-multiplicity basis of a block's SU(2) CGT space: array of shape
    (om, 2j_1+1, ..., 2j_r+1), each basis tensor of unit norm (index i = j - m
    per leg); om = 0 when the labels admit no invariant."""
    L = _capi.lib()
    P = C.POINTER
    t, d = _ivec(tj), _ivec(dirs)
    err = C.create_string_buffer(256)
    om = L.qsx_su2_om_basis(len(tj), t.ctypes.data_as(P(C.c_int32)), d.ctypes.data_as(P(C.c_int32)), None, 0, err, 256)
    if om < 0:
        _capi.check(-om, err)
    shape = tuple(int(x) + 1 for x in tj)
    n = int(np.prod(shape, dtype=np.int64))
    out = np.zeros(om * n)
    if om:
        rc = L.qsx_su2_om_basis(len(tj), t.ctypes.data_as(P(C.c_int32)), d.ctypes.data_as(P(C.c_int32)),
                                out.ctypes.data_as(P(C.c_double)), out.size, err, 256)
        if rc < 0:
            _capi.check(-rc, err)
    return np.stack([out[a * n:(a + 1) * n].reshape(shape, order="F") for a in range(om)]) if om else \
        np.zeros((0,) + shape)


def su2_canon_order(tj, dirs) -> list:
    """canonical leg order of a block CGT: ord[k] = block leg at position k"""
    t, d = _ivec(tj), _ivec(dirs)
    o = np.zeros(max(1, len(tj)), np.int32)
    P = C.POINTER
    rc = _capi.lib().qsx_su2_canon_order(len(tj), t.ctypes.data_as(P(C.c_int32)), d.ctypes.data_as(P(C.c_int32)),
                                         o.ctypes.data_as(P(C.c_int32)))
    _capi.check(rc, None)
    return [int(x) for x in o[:len(tj)]]


def su2_xom(tja, dira, tjb, dirb, ia, ib, perm=None) -> np.ndarray:
    """matrix-valued X-symbol x[a, b, c] = <C_c | contract(A_a, B_b)> of the
    blocks' outer-multiplicity bases (output legs [kept A | kept B] permuted
    by perm)."""
    L = _capi.lib()
    P = C.POINTER
    rc_ = len(tja) + len(tjb) - 2 * len(ia)
    a, da, b, db, ca, cb = (_ivec(v) for v in (tja, dira, tjb, dirb, ia, ib))
    pv = _ivec(perm if perm is not None else range(rc_))
    dims = (C.c_int32 * 3)()
    err = C.create_string_buffer(256)
    args = (len(tja), a.ctypes.data_as(P(C.c_int32)), da.ctypes.data_as(P(C.c_int32)), len(tjb),
            b.ctypes.data_as(P(C.c_int32)), db.ctypes.data_as(P(C.c_int32)), ca.ctypes.data_as(P(C.c_int32)),
            cb.ctypes.data_as(P(C.c_int32)), len(ia), pv.ctypes.data_as(P(C.c_int32)), dims)
    _capi.check(L.qsx_su2_xom(*args, None, 0, err, 256), err)
    x = np.zeros(dims[0] * dims[1] * dims[2])
    _capi.check(L.qsx_su2_xom(*args, x.ctypes.data_as(P(C.c_double)), x.size, err, 256), err)
    return x.reshape(dims[0], dims[1], dims[2])
