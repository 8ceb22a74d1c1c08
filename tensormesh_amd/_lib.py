"""ctypes binding of ``libtensormesh_amd.so`` (the C ABI declared in
``include/tensormesh_amd.h``).

torch is imported before the library is opened so that the library binds to
the HIP runtime torch already loaded (same ``libamdhip64.so.7`` soname).  A
missing library raises immediately: there is no alternative code path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtensormesh_amd.so")

TM_OK, TM_ERR_ARG, TM_ERR_HIP, TM_ERR_LIMIT, TM_ERR_ALLOC, TM_ERR_NOTCONV = 0, -1, -2, -3, -4, -5
FORM_LAPLACE, FORM_KAPPA_LAPLACE, FORM_MASS = 0, 1, 2
FORM_GENERAL = 3   # Python-side tag: dispatched to tm_asm_p1_general
LOAD_CONST, LOAD_NODAL = 0, 1
ASM_DETERMINISTIC = 0x100   # TM_ASM_DETERMINISTIC
F32, F64, BF16 = 0, 1, 2

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_double = ctypes.c_double


class PlanDev(ctypes.Structure):
    _fields_ = [("blk", c_void_p), ("halo", c_void_p), ("meta", c_void_p), ("recs", c_void_p),
                ("rec_elem", c_void_p), ("runs", c_void_p), ("n_blocks", ctypes.c_int32),
                ("lds_bytes", ctypes.c_int32), ("threads", ctypes.c_int32), ("entry_bytes", ctypes.c_int32),
                ("erecs", c_void_p), ("eblk", c_void_p), ("emir", c_void_p), ("erows", c_void_p),
                ("erec_words", ctypes.c_int32), ("elds_bytes", ctypes.c_int32)]


class Quad(ctypes.Structure):
    _fields_ = [("wsum", c_double), ("W", c_double * 4), ("M", c_double * 16)]


class BatchedArgs(ctypes.Structure):
    _fields_ = [("n_rows", c_i64), ("n_tiles", ctypes.c_int32), ("n_mesh", ctypes.c_int32),
                ("row_ptr", c_void_p), ("col_idx", c_void_p), ("values", c_void_p),
                ("value_dtype", ctypes.c_int32), ("reserved", ctypes.c_int32), ("b", c_void_p),
                ("x", c_void_p), ("tile_row0", c_void_p), ("tile_mesh", c_void_p),
                ("mesh_tile0", c_void_p), ("rtol", c_double), ("maxiter", c_i64),
                ("work", c_void_p), ("mesh_status", c_void_p), ("mesh_iters", c_void_p),
                ("mesh_relres", c_void_p)]


class FcgView(ctypes.Structure):
    _fields_ = [("r", c_void_p), ("z", c_void_p), ("ss_part", c_void_p), ("scale", c_void_p),
                ("relres", c_void_p), ("status", c_void_p), ("iter", c_void_p), ("nparts", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("dr", c_void_p)]


class TMError(RuntimeError):
    """Failure reported by the native library."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[tensormesh_amd error {code}] {msg}")
        self.code = code


class NotConverged(TMError):
    pass


_SIGS = {
    "tm_last_error": (ctypes.c_char_p, []),
    "tm_version": (c_int, []),
    "tm_pattern_rowptr": (c_int, [c_void_p, c_i64, c_int, c_i64, c_void_p]),
    "tm_pattern_colidx": (c_int, [c_void_p, c_i64, c_int, c_i64, c_void_p, c_void_p]),
    "tm_element_slots": (c_int, [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p]),
    "tm_plan_build": (c_int, [c_void_p, c_int, c_void_p, c_i64, c_int, c_i64, c_void_p, c_void_p,
                              c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_void_p)]),
    "tm_plan_info": (c_int, [c_void_p, c_void_p]),
    "tm_plan_export": (c_int, [c_void_p] * 7),
    "tm_plan_destroy": (None, [c_void_p]),
    "tm_plan_edge_build": (c_int, [c_int, c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "tm_plan_edge_info": (c_int, [c_void_p, c_void_p]),
    "tm_plan_edge_export": (c_int, [c_void_p] * 5),
    "tm_plan_edge_destroy": (None, [c_void_p]),
    "tm_asm_p1_matrix": (c_int, [c_int, c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_int,
                                 c_double, ctypes.POINTER(Quad), c_void_p, c_int, c_void_p]),
    "tm_asm_p1_general": (c_int, [c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p,
                                  ctypes.POINTER(Quad), c_void_p, c_int, c_int, c_void_p]),
    "tm_asm_p1_load": (c_int, [c_int, c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_int,
                               c_double, ctypes.POINTER(Quad), c_void_p, c_void_p]),
    "tm_asm_p1_kappa_adjoint": (c_int, [c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_int,
                                        ctypes.POINTER(Quad), c_void_p, c_void_p]),
    "tm_spmv": (c_int, [c_i64, c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "tm_gather": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "tm_pattern_dev_work_size": (c_i64, [c_i64, c_int, c_i64]),
    "tm_pattern_dev_rows": (c_int, [c_void_p, c_i64, c_int, c_i64, c_void_p, c_void_p, c_void_p]),
    "tm_pattern_dev_cols": (c_int, [c_void_p, c_i64, c_int, c_i64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_element_slots_dev": (c_int, [c_void_p, c_i64, c_int, c_void_p, c_void_p, c_void_p, c_i64, c_void_p,
                                     c_void_p]),
    "tm_condense_work_size": (c_i64, [c_i64]),
    "tm_condense_count": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_condense_fill": (c_int, [c_i64] + [c_void_p] * 17 + [c_void_p]),
    "tm_condense_rhs": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                c_void_p, c_void_p]),
    "tm_condense_recover": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_spmv_bsr3": (c_int, [c_i64, c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "tm_pcg_jacobi_bsr3": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_double, c_double, c_i64, c_int, ctypes.POINTER(c_i64),
                                   ctypes.POINTER(c_double), c_void_p]),
    "tm_asm_nh": (c_int, [ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_void_p, c_int, c_double, c_double,
                          c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "tm_asm_linear_elastic": (c_int, [ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_int, c_double, c_double,
                                      c_void_p, c_int, c_void_p]),
    "tm_nh_apply": (c_int, [ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_void_p, c_int, c_double, c_double,
                            c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_batched_pcg_work_size": (c_i64, [c_i64, ctypes.c_int32]),
    "tm_batched_pcg": (c_int, [ctypes.POINTER(BatchedArgs), c_int, c_int, c_void_p]),
    "tm_batched_timeout_flag": (c_int, [ctypes.POINTER(BatchedArgs), c_void_p, c_void_p]),
    "tm_gnn_mp_forward": (c_int, [c_i64] + [c_void_p] * 6 + [c_int] + [c_void_p] * 5),
    "tm_gnn_part_floats": (c_i64, []),
    "tm_gnn_mp_backward": (c_int, [c_i64] + [c_void_p] * 8 + [c_int] + [c_void_p] * 5),
    "tm_gnn_pool": (c_int, [c_i64] + [c_void_p] * 5 + [c_int, c_void_p, c_void_p]),
    "tm_gnn_unpool": (c_int, [c_i64] + [c_void_p] * 5 + [c_int, c_void_p, c_void_p]),
    "tm_gnn_sumsq": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_int), c_void_p]),
    "tm_gnn_scale": (c_int, [c_i64, c_void_p, c_int, c_void_p, c_void_p]),
    "tm_gnn_encode": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_int] + [c_void_p] * 6),
    "tm_gnn_encode_backward": (c_int, [c_i64] + [c_void_p] * 8),
    "tm_gnn_decode": (c_int, [c_i64] + [c_void_p] * 9),
    "tm_gnn_encode_pool": (c_int, [c_i64, c_i64] + [c_void_p] * 7 + [c_int] + [c_void_p] * 7),
    "tm_gnn_final": (c_int, [c_i64] + [c_void_p] * 15 + [c_int] + [c_void_p] * 7),
    "tm_gnn_level_prolong": (c_int, [c_i64] + [c_void_p] * 7),
    "tm_gnn_unpool_decode": (c_int, [c_i64, c_i64] + [c_void_p] * 14),
    "tm_gnn_sym0_down": (c_int, [c_i64] + [c_void_p] * 6 + [c_int] + [c_void_p] * 4),
    "tm_gnn_sym0_up": (c_int, [c_i64, c_i64] + [c_void_p] * 18),
    "tm_gnn_row_sums": (c_int, [c_i64] + [c_void_p] * 4),
    "tm_gnn_decode_backward": (c_int, [c_i64] + [c_void_p] * 7),
    "tm_gnn_energy_loss": (c_int, [c_i64] + [c_void_p] * 10 + [ctypes.POINTER(c_int), c_void_p]),
    "tm_gnn_level_setup": (c_int, [c_i64] + [c_void_p] * 6),
    "tm_gnn_cosine_grad": (c_int, [c_i64] + [c_void_p] * 4 + [c_int, c_void_p, c_void_p]),
    "tm_gnn_level_coarsen": (c_int, [c_i64] + [c_void_p] * 5),
    "tm_gnn_pool_weights": (c_int, [c_i64] + [c_void_p] * 5),
    "tm_gnn_adam": (c_int, [c_i64] + [c_void_p] * 5 + [ctypes.c_float] * 4 + [c_i64, c_void_p]),
    "tm_fcg_work_size": (c_i64, [c_i64, c_int]),
    "tm_fcg_layout": (c_int, [c_i64, c_int, c_void_p, ctypes.POINTER(FcgView)]),
    "tm_fcg_start": (c_int, [c_i64, c_int] + [c_void_p] * 7 + [c_double, c_double, c_i64, c_void_p]),
    "tm_fcg_iterate": (c_int, [c_i64, c_int] + [c_void_p] * 7),
    "tm_pcg_work_size": (c_i64, [c_i64]),
    "tm_bicgstab_work_size": (c_i64, [c_i64]),
    "tm_bicgstab_jacobi": (c_int, [c_i64, c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                   c_double, c_double, c_i64, c_int, ctypes.POINTER(c_i64),
                                   ctypes.POINTER(c_double), c_void_p]),
    "tm_pcg_jacobi": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                              c_double, c_double, c_i64, c_int, ctypes.POINTER(c_i64),
                              ctypes.POINTER(c_double), c_void_p]),
    "tm_csr_delta16": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "tm_pcg_jacobi_d16": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                  c_double, c_double, c_i64, c_int, ctypes.POINTER(c_i64),
                                  ctypes.POINTER(c_double), c_void_p]),
    "tm_asm_p1_moments": (c_int, [c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_void_p, c_int,
                                  c_int, c_void_p]),
    "tm_asm_p1_load_moments": (c_int, [c_int, c_int, ctypes.POINTER(PlanDev), c_void_p, c_void_p, c_void_p,
                                       c_int, c_void_p]),
    "tm_facet_load": (c_int, [c_int, c_int, c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int,
                              c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "tm_pcg_noalloc_max_rows": (c_i64, []),
    "tm_pcg_jacobi_noalloc": (c_int, [c_i64, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                      c_double, c_double, c_i64, c_int, ctypes.POINTER(c_i64),
                                      ctypes.POINTER(c_double), c_void_p]),
}

_lib = None


def lib():
    """The loaded native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              f"`python -m tensormesh_amd.csrc.build`")
        h = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def declared_symbols():
    return list(_SIGS)


def check(status: int):
    if status == TM_OK:
        return
    msg = lib().tm_last_error().decode(errors="replace")
    if status == TM_ERR_NOTCONV:
        raise NotConverged(status, msg)
    if status == TM_ERR_ARG:
        raise ValueError(msg)
    raise TMError(status, msg)


def ptr(t) -> c_void_p:
    if t is None:
        return c_void_p(None)
    if isinstance(t, np.ndarray):
        return c_void_p(t.ctypes.data)
    return c_void_p(t.data_ptr())


def stream_of(device: torch.device) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return F32
    if t.dtype == torch.float64:
        return F64
    raise TypeError(f"unsupported dtype {t.dtype} (float32 or float64 expected)")


def require_gpu(t: torch.Tensor, what: str):
    if not t.is_cuda:
        raise RuntimeError(f"{what} must be on a GPU device for the HIP path, got {t.device}")
