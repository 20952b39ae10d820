"""ctypes binding of librespack_gpu.so (declared in include/respack_gpu.h).

There is no fallback: if the library is missing, or no gfx950 device is
visible, the calls raise.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librespack_gpu.so")

P = C.c_void_p
I32 = C.c_int32
D = C.c_double

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}

# Every symbol include/respack_gpu.h declares.
EXPORTS = [
    "rp_init", "rp_finalize", "rp_last_error", "rp_algn235", "rp_gen_grid", "rp_make_kg0",
    "rp_ttrhdrn_mkidx", "rp_chiqw_open", "rp_chiqw_set_crpa", "rp_chiqw_info", "rp_chiqw_kmesh",
    "rp_chiqw_prepare_q", "rp_pairdens_batch", "rp_tetra_weights_batch", "rp_chi_accumulate",
    "rp_dyson", "rp_dyson_noalloc", "rp_chiqw_q", "rp_chiqw_npair", "rp_chiqw_set_window",
    "rp_mlp_open", "rp_mlp_nparams", "rp_mlp_get_params", "rp_mlp_set_params", "rp_mlp_forward",
    "rp_mlp_gradient", "rp_mlp_train", "rp_chi_set",
    "rp_w3d_open", "rp_w3d_info", "rp_w3d_kmesh", "rp_w3d_set_wannier", "rp_w3d_set_q",
    "rp_w3d_set_epsinv", "rp_w3d_funcs", "rp_w3d_vw", "rp_j3d_run", "rp_w3d_stage_times",
    "rp_chiqw_batch_plan", "rp_chiqw_batch_run", "rp_chiqw_batch_result", "rp_chiqw_batch_free",
    "rp_chiqw_autotune", "rp_chiqw_tune_count", "rp_chiqw_tune_info", "rp_chiqw_set_span",
]


class RpError(RuntimeError):
    pass


class ChiqwParams(C.Structure):
    _fields_ = [("n_calc_band", I32), ("num_freq_grid", I32), ("flg_crpa", I32), ("precision", I32),
                ("ecut_for_eps", D), ("shift_ef", D), ("max_excitation_energy", D),
                ("green_func_delt", D), ("ttrhdrn_dmna", D), ("ttrhdrn_dmnr", D)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RpError(f"{LIB_PATH} is not built (python -m respackdev_amd.build)")
        L = C.CDLL(LIB_PATH)
        L.rp_last_error.restype = C.c_char_p
        L.rp_init.argtypes = [I32, C.POINTER(P)]
        L.rp_finalize.argtypes = [P]
        L.rp_algn235.restype = I32
        L.rp_algn235.argtypes = [I32]
        L.rp_gen_grid.argtypes = [I32, D, P]
        L.rp_make_kg0.argtypes = [I32, P, D, P, P, P]
        L.rp_ttrhdrn_mkidx.argtypes = [I32, I32, I32, P, P]
        L.rp_chiqw_open.argtypes = [P, P, I32, I32, P, P, I32, P, P, I32, P, I32, P, P, D,
                                    C.POINTER(ChiqwParams)]
        L.rp_chiqw_set_crpa.argtypes = [P, I32, P, P, I32, I32, P]
        L.rp_chiqw_info.argtypes = [P, P, P, P]
        L.rp_chiqw_kmesh.argtypes = [P, P, P, P]
        L.rp_chiqw_prepare_q.argtypes = [P, P, P, P, I32]
        L.rp_pairdens_batch.argtypes = [P, P]
        L.rp_tetra_weights_batch.argtypes = [P, P]
        L.rp_chi_accumulate.argtypes = [P, I32, P]
        L.rp_dyson.argtypes = [P, P]
        L.rp_dyson_noalloc.argtypes = [P, P, P]
        L.rp_chiqw_q.argtypes = [P, P, I32, P, P, P]
        L.rp_chiqw_npair.argtypes = [P, P]
        L.rp_chiqw_set_window.argtypes = [P, D, I32]
        L.rp_chiqw_batch_plan.argtypes = [P, I32, P, P, P, I32]
        L.rp_chiqw_batch_run.argtypes = [P, P]
        L.rp_chiqw_batch_result.argtypes = [P, I32, P, P, P, I32, P]
        L.rp_chiqw_batch_free.argtypes = [P]
        L.rp_chiqw_autotune.argtypes = [P, I32]
        L.rp_chiqw_set_span.argtypes = [P, I32]
        L.rp_chiqw_tune_count.argtypes = [P, P]
        L.rp_chiqw_tune_info.argtypes = [P, I32, P, P]
        I64, U64 = C.c_int64, C.c_uint64
        L.rp_mlp_open.argtypes = [P, I32, I32, I32, I32, U64]
        L.rp_mlp_nparams.argtypes = [P, P]
        L.rp_mlp_get_params.argtypes = [P, P]
        L.rp_mlp_set_params.argtypes = [P, P]
        L.rp_mlp_forward.argtypes = [P, I64, P, P]
        L.rp_mlp_gradient.argtypes = [P, I64, P, P, P, P]
        L.rp_mlp_train.argtypes = [P, I64, P, P, I32, I32, D, U64, P]
        L.rp_chi_set.argtypes = [P, P]
        L.rp_w3d_open.argtypes = [P, P, I32, I32, P, P, I32, P, P, I32, P, D, I32, P]
        L.rp_w3d_info.argtypes = [P, P, P]
        L.rp_w3d_kmesh.argtypes = [P, P, P]
        L.rp_w3d_set_wannier.argtypes = [P, I32, P]
        L.rp_w3d_set_q.argtypes = [P, I32, P, P]
        L.rp_w3d_set_epsinv.argtypes = [P, I32, P]
        L.rp_w3d_funcs.argtypes = [P, P, P]
        L.rp_w3d_vw.argtypes = [P, P, P, P]
        L.rp_j3d_run.argtypes = [P, I32, P, P, P, P]
        L.rp_w3d_stage_times.argtypes = [P, P]
        for name in EXPORTS:
            getattr(L, name).restype = getattr(L, name).restype or I32
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise RpError(lib().rp_last_error().decode())


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
