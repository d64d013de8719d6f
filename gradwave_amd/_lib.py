"""ctypes binding of libgradwave_hip.so (declarations in include/gradwave_hip.h).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises. There is no CPU or PyTorch fallback.
"""

from __future__ import annotations

import ctypes as C
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgradwave_hip.so")

GW_FP64, GW_FP32 = 0, 1
GW_TERM_KINETIC, GW_TERM_LOCAL, GW_TERM_NONLOCAL, GW_TERM_ALL = 1, 2, 4, 7
GW_PLAN_FORCE_UNFUSED, GW_PLAN_TIMING, GW_PLAN_LEGACY, GW_PLAN_PENCIL = 1, 2, 4, 8
GW_TUNE_RADIX, GW_TUNE_GEMM_WGS, GW_TUNE_GEMM_TILE, GW_TUNE_XPASS, GW_TUNE_COMBINE_LIB, GW_TUNE_INNER_LIB = 1, 2, 3, 4, 5, 6
GW_TUNE_BAND_CHUNK = 7
GW_TUNE_PLANE_BANDS = 8
STAGES = ("plane_inv", "xpass", "plane_fwd", "becp", "dmul", "backproj")

_P = C.c_void_p
_I = C.c_int
_L = C.c_long
_Z = C.c_size_t
_D = C.c_double

# name -> (restype, argtypes); mirrors include/gradwave_hip.h
SIGNATURES = {
    "gw_last_error": (C.c_char_p, []),
    "gw_version": (_I, []),
    "gw_plan_create": (_I, [C.POINTER(_P), _I, _I, _I, _I, _I, _P, _P]),
    "gw_plan_destroy": (_I, [_P]),
    "gw_box_split": (_I, [_I, _P]),
    "gw_plan_set_flags": (_I, [_P, _I]),
    "gw_plan_stage_times": (_I, [_P, _P, _P, _I, _I]),
    "gw_plan_info": (_I, [_P, _I, _P, _I]),
    "gw_plan_set_kinetic": (_I, [_P, _P, _P]),
    "gw_plan_set_potential": (_I, [_P, _P, _P]),
    "gw_plan_set_projectors": (_I, [_P, _I, _P, _P, _P]),
    "gw_hpsi_workspace_bytes": (_Z, [_P, _I, _I]),
    "gw_hpsi_apply": (_I, [_P, _I, _I, _P, _P, _I, _P, _Z, _P]),
    "gw_hpsi_apply_strided": (_I, [_P, _I, _I, _P, _L, _L, _P, _L, _L, _I, _P, _Z, _P]),
    "gw_hpsi_noalloc_max_bytes": (_Z, []),
    "gw_hpsi_apply_noalloc": (_I, [_P, _I, _I, _P, _L, _L, _P, _L, _L, _I, _P]),
    "gw_wfc_to_real_workspace_bytes": (_Z, [_P, _I, _I]),
    "gw_wfc_to_real": (_I, [_P, _I, _P, _P, _I, _P, _Z, _P]),
    "gw_sphere_to_box": (_I, [_P, _I, _P, _P, _I, _P]),
    "gw_box_to_sphere": (_I, [_P, _I, _P, _P, _I, _P]),
    "gw_zgemm_inner_workspace_bytes": (_Z, [_I, _I, _I, _I, _I]),
    "gw_zgemm_inner": (_I, [_I, _I, _I, _I, _I, _P, _L, _L, _P, _L, _L, _P, _L, _L, _P, _Z, _P]),
    "gw_zgemm_combine": (_I, [_I, _I, _I, _I, _I, _P, _L, _L, _I, _P, _L, _L, _P, _L, _L, _D, _P]),
    "gw_residual": (_I, [_I, _I, _I, _I, _P, _P, _L, _L, _P, _P, _P, _P]),
    "gw_teter_expand": (_I, [_I, _I, _I, _I, _I, _P, _P, _L, _L, _P, _L, _P, _P, _L, _L, _P]),
    "gw_row_norms": (_I, [_I, _I, _I, _I, _P, _L, _L, _P, _P]),
    "gw_heev_workspace_bytes": (_Z, [_I, _I]),
    "gw_heev_max_n": (_I, []),
    "gw_heev_batched": (_I, [_I, _P, _P]),
    "gw_chol_max_n": (_I, []),
    "gw_chol_inv_batched": (_I, [_I, _P, _P]),
    "gw_cg_update": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _L, _P, _L, _P, _P, _P, _P, _P]),
    "gw_cg_direction": (_I, [_I, _I, _I, _P, _P, _P, _P, _L, _P, _L, _P]),
    "gw_row_dots": (_I, [_I, _I, _I, _P, _P, _P, _L, _P, _L, _P, _P, _P]),
    "gw_hpsi_adjoint_v_workspace_bytes": (_Z, [_P, _I, _I]),
    "gw_hpsi_adjoint_v": (_I, [_P, _I, _P, _L, _L, _P, _L, _L, _I, _P, _P, _Z, _P]),
    "gw_density_workspace_bytes": (_Z, [_P, _I, _I]),
    "gw_density": (_I, [_P, _I, _P, _L, _L, _I, _P, _P, _P, _Z, _P]),
    "gw_proj_phase": (_I, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "gw_proj_phase_grad_workspace_bytes": (_Z, [_I, _I]),
    "gw_proj_phase_grad": (_I, [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _Z, _P]),
    "gw_vloc_g": (_I, [_I, _L, _P, _P, _P, _P, _D, _P, _P]),
    "gw_vloc_g_grad_workspace_bytes": (_Z, [_I]),
    "gw_vloc_g_grad": (_I, [_I, _L, _P, _P, _P, _P, _D, _P, _P, _P, _Z, _P]),
    "gw_rgemm_inner_workspace_bytes": (_Z, [_I, _I, _I, _I, _I]),
    "gw_rgemm_inner": (_I, [_I, _I, _I, _I, _I, _P, _L, _L, _P, _L, _L, _P, _L, _L, _P, _Z, _P]),
    "gw_rgemm_combine": (_I, [_I, _I, _I, _I, _I, _P, _L, _L, _P, _L, _L, _P, _L, _L, _D, _D, _P]),
    "gw_tune_set": (_I, [_I, _I, _I]),
    "gw_tune_get": (_I, [_I, _I]),
    "gw_batch_create": (_I, [C.POINTER(_P), _P, _I]),
    "gw_batch_destroy": (_I, [_P]),
    "gw_batch_workspace_bytes": (_Z, [_P, _P, _I]),
    "gw_batch_bind": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _Z, _P, C.POINTER(_P)]),
    "gw_batch_apply_bound": (_I, [_P, _P]),
    "gw_batch_binding_destroy": (_I, [_P]),
    "gw_batch_apply": (_I, [_P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _Z, _P]),
    "gw_toeplitz_build": (_I, [_I, _I, _I, _I, _I, _I, _P, _P, _P, _P, _P]),
    "gw_gamma_pack": (_I, [_I, _I, _I, _I, _P, _L, _P, _P, _P, _L, _P]),
    "gw_gamma_unpack": (_I, [_I, _I, _I, _P, _L, _P, _P, _P, _L, _P]),
}



class HeevJob(C.Structure):
    """gw_heev_job of include/gradwave_hip.h."""
    _fields_ = [("a", _P), ("w", _P), ("z", _P), ("ws", _P),
                ("n", _I), ("lda", _I), ("nev", _I), ("ldz", _I)]


class CholJob(C.Structure):
    """gw_chol_job of include/gradwave_hip.h."""
    _fields_ = [("a", _P), ("linv", _P), ("info", _P), ("n", _I), ("lda", _I), ("ldl", _I)]


_lib = None


class GWError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """The loaded library; raises GWError if it is not built or fails to load."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GWError(f"{LIB_PATH} not built: run python -m gradwave_amd.build")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
        from gradwave_amd import autotune  # saved radix / split-K choices, if any

        autotune.load()
    return _lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().gw_last_error()
        raise GWError(f"{what} failed (code {rc}): {msg.decode() if msg else ''}")


def stream_ptr(device: torch.device | None = None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()


def precision_of(dtype: torch.dtype) -> int:
    if dtype in (torch.complex128, torch.float64):
        return GW_FP64
    if dtype in (torch.complex64, torch.float32):
        return GW_FP32
    raise TypeError(f"unsupported dtype {dtype}")


def require_gpu(t: torch.Tensor, name: str) -> None:
    if not t.is_cuda:
        raise GWError(f"{name} must be a GPU tensor (gradwave_amd has no CPU path)")
