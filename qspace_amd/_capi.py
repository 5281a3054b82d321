"""ctypes binding of the in-tree libqsx.so (the HIP implementation).

There is no fallback: if the library is missing or cannot be loaded the
import raises, so no caller can silently run without the native code.
"""
from __future__ import annotations

import ctypes as C
import os

from .qspace import QsxCell, QsxOpts, QsxTensor, QsxTensorOut


class QsxOrthoInfo(C.Structure):
    """qsx_ortho_info (include/qsx.h)"""
    _fields_ = [("ntot", C.c_int64), ("nkeep", C.c_int64), ("nkmin", C.c_int32), ("flag", C.c_int32),
                ("stol", C.c_double), ("svd2tr", C.c_double), ("sv2tot", C.c_double), ("sfac", C.c_double),
                ("svd", C.POINTER(C.c_double)), ("gpu_ms", C.c_double), ("sweeps", C.c_int32),
                ("ngroups", C.c_int32), ("svd_qdim", C.POINTER(C.c_double))]

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libqsx.so")

_lib = None


class QsxError(RuntimeError):
    """Error returned by the qsx C-ABI (code = QSX_ERR_*)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[qsx error {code}] {msg}")
        self.code = code


ERR = {1: "ARG", 2: "ITAG", 3: "QDIM", 4: "STRUCT", 5: "UNSUPPORTED", 6: "DEVICE", 7: "NOMEM", 8: "CG"}
MODES = {"batched": 0, "plain": 1, "split_bf16": 2, "split_f32": 3}


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ImportError(f"qsx native library not built: {LIBPATH} (run __graft_entry__.build())")
    L = C.CDLL(LIBPATH)
    P = C.POINTER
    i32, i64, f64, vp, sz = C.c_int32, C.c_int64, C.c_double, C.c_void_p, C.c_size_t
    cs = C.c_char_p
    sig = {
        "qsx_version": ([], cs),
        "qsx_device_init": ([C.c_int, cs, sz], C.c_int),
        "qsx_malloc_device": ([P(vp), sz, cs, sz], C.c_int),
        "qsx_free_device": ([vp], C.c_int),
        "qsx_memcpy_h2d": ([vp, vp, sz, vp], C.c_int),
        "qsx_memcpy_d2h": ([vp, vp, sz, vp], C.c_int),
        "qsx_memset_device": ([vp, C.c_int, sz, vp], C.c_int),
        "qsx_synchronize": ([], C.c_int),
        "qsx_copy_device": ([vp, vp, i64, vp], C.c_int),
        "qsx_contract": ([P(QsxTensor), P(i32), i32, i32, P(QsxTensor), P(i32), i32, i32, P(i32), i32, P(QsxOpts),
                          P(P(QsxTensorOut)), cs, sz], C.c_int),
        "qsx_tensor_free": ([P(QsxTensorOut)], None),
        "qsx_contract_cell": ([P(QsxCell), i32, P(i32), i32, P(QsxOpts), P(P(QsxTensorOut)), cs, sz], C.c_int),
        "qsx_itag_match": ([P(QsxTensor), i32, P(C.c_char_p), P(i32), i32, P(QsxTensor), i32, P(C.c_char_p), P(i32),
                            i32, P(i32), P(i32), P(i32), P(i32), P(i32), cs, sz], C.c_int),
        "qsx_permute": ([P(QsxTensor), P(i32), i32, i32, P(P(QsxTensorOut)), cs, sz], C.c_int),
        "qsx_plus": ([P(QsxTensor), P(QsxTensor), f64, P(QsxOpts), P(P(QsxTensorOut)), cs, sz], C.c_int),
        "qsx_ortho": ([P(QsxTensor), i32, i32, i64, f64, cs, P(P(QsxTensorOut)), P(P(QsxTensorOut)),
                       P(P(QsxOrthoInfo)), cs, sz], C.c_int),
        "qsx_ortho_info_free": ([P(QsxOrthoInfo)], None),
        "qsx_parse_ctridx": ([cs, P(i32), i32, P(i32)], C.c_int),
        "qsx_packed_offsets": ([P(QsxTensor), P(i64)], C.c_int),
        "qsx_ctr_plan_create": ([P(QsxTensor), P(i32), i32, i32, P(QsxTensor), P(i32), i32, i32, P(i32), i32,
                                 P(QsxOpts), P(vp), cs, sz], C.c_int),
        "qsx_ctr_plan_output": ([vp, P(P(QsxTensorOut))], C.c_int),
        "qsx_ctr_plan_pairs": ([vp, P(i64), P(i64), P(i32), P(i32), P(f64), P(i64)], C.c_int),
        "qsx_ctr_plan_idxset": ([vp, P(i64), P(i64), P(i32), P(i32), P(f64), P(i64), P(i32)], C.c_int),
        "qsx_ctr_execute": ([vp, vp, vp, vp, vp], C.c_int),
        "qsx_ctr_plan_flops": ([vp], f64),
        "qsx_block_zero_flags": ([P(i64), i32, vp, P(i32), cs, sz], C.c_int),
        "qsx_block_zero_flags_noalloc": ([P(i64), i32, vp, P(i32), cs, sz], C.c_int),
        "qsx_ctr_plan_set_policy": ([vp, P(i32)], C.c_int),
        "qsx_hpsi_set_policy": ([vp, P(i32)], C.c_int),
        "qsx_ctr_plan_destroy": ([vp], None),
        "qsx_hpsi_plan_create": ([P(QsxTensor), P(QsxTensor), P(QsxTensor), P(QsxOpts), P(vp), cs, sz], C.c_int),
        "qsx_hpsi_plan_create_pseudo": ([P(QsxTensor), P(QsxTensor), i32, P(QsxTensor), P(QsxTensor), P(i32),
                                         P(QsxTensor), P(QsxOpts), P(vp), cs, sz], C.c_int),
        "qsx_hpsi_size": ([vp], i64),
        "qsx_hpsi_flops": ([vp], f64),
        "qsx_hpsi_bytes": ([vp], f64),
        "qsx_hpsi_apply": ([vp, vp, vp, vp], C.c_int),
        "qsx_hpsi_use_graph": ([vp, i32], C.c_int),
        "qsx_hpsi_weights": ([vp, P(f64)], C.c_int),
        "qsx_hpsi_stage_info": ([vp, P(f64)], C.c_int),
        "qsx_hpsi_autotune": ([vp, i32, P(f64), P(i32)], C.c_int),
        "qsx_hpsi_last_timing": ([vp, P(f64)], C.c_int),
        "qsx_hpsi_stage_ms": ([vp, i32, i32, vp, vp, P(f64)], C.c_int),
        "qsx_hpsi_plan_destroy": ([vp], None),
        "qsx_lanczos": ([vp, P(f64), i32, P(f64), P(f64), P(f64), cs, sz], C.c_int),
        "qsx_davidson": ([vp, P(f64), i32, P(f64), P(f64), P(f64), cs, sz], C.c_int),
        "qsx_su2_cg": ([i32] * 6, f64),
        "qsx_su2_x3": ([i32, P(i32), P(i32), i32, P(i32), P(i32), P(i32), P(i32), i32, P(f64), cs, sz], C.c_int),
        "qsx_su2_om_basis": ([i32, P(i32), P(i32), P(f64), i64, cs, sz], C.c_int),
        "qsx_su2_canon_order": ([i32, P(i32), P(i32), P(i32)], C.c_int),
        "qsx_su2_xom": ([i32, P(i32), P(i32), i32, P(i32), P(i32), P(i32), P(i32), i32, P(i32), P(i32), P(f64), i64,
                         cs, sz], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def exported_symbols() -> list[str]:
    """names declared in include/qsx.h that the loaded library must export."""
    return ["qsx_version", "qsx_device_init", "qsx_malloc_device", "qsx_free_device", "qsx_memcpy_h2d",
            "qsx_memcpy_d2h", "qsx_memset_device", "qsx_synchronize", "qsx_copy_device", "qsx_contract", "qsx_tensor_free", "qsx_contract_cell", "qsx_itag_match",
            "qsx_permute", "qsx_plus", "qsx_ortho", "qsx_ortho_info_free",
            "qsx_parse_ctridx", "qsx_packed_offsets", "qsx_ctr_plan_create", "qsx_ctr_plan_output",
            "qsx_ctr_plan_pairs", "qsx_ctr_plan_idxset", "qsx_ctr_execute", "qsx_ctr_plan_flops", "qsx_ctr_plan_set_policy", "qsx_ctr_plan_destroy", "qsx_block_zero_flags", "qsx_block_zero_flags_noalloc",
            "qsx_hpsi_plan_create", "qsx_hpsi_plan_create_pseudo", "qsx_hpsi_size", "qsx_hpsi_flops", "qsx_hpsi_bytes", "qsx_hpsi_apply",
            "qsx_hpsi_use_graph", "qsx_hpsi_weights", "qsx_hpsi_stage_info", "qsx_hpsi_autotune", "qsx_hpsi_set_policy", "qsx_hpsi_last_timing", "qsx_hpsi_stage_ms", "qsx_hpsi_plan_destroy", "qsx_lanczos", "qsx_davidson",
            "qsx_su2_cg", "qsx_su2_x3", "qsx_su2_om_basis", "qsx_su2_xom",
            "qsx_su2_canon_order"]


def check(rc: int, err) -> None:
    if rc:
        msg = err.value.decode(errors="replace") if err is not None else ""
        raise QsxError(rc, msg)
