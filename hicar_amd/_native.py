"""Loader of libwindhip.so (the HIP solver) and ctypes mirrors of include/windhip.h.

The library is built in-tree by ``__graft_entry__.build()`` (hicar_amd/csrc/Makefile).
There is no fallback: if the library is missing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libwindhip.so")

WINDHIP_OK = 0
WINDHIP_ERR_ARG = 1
WINDHIP_ERR_HIP = 2
WINDHIP_ERR_NAN = 3
WINDHIP_ERR_DIVERGED = 4
WINDHIP_ERR_NOMEM = 5

WINDHIP_F32 = 4
WINDHIP_F64 = 8
WINDHIP_HOST = 0
WINDHIP_DEVICE = 1

START_WARM, START_ZERO, START_GIVEN, START_LEARNED = 0, 1, 2, 3
PLAIN_F64, MIXED, MIXED_LINE, MIXED_MG = 0, 1, 2, 3

# every entry point declared in include/windhip.h
EXPORTS = (
    "windhip_default_opts", "windhip_create", "windhip_destroy", "windhip_last_error",
    "windhip_set_metrics", "windhip_solve", "windhip_apply_correction", "windhip_calc_iter_winds",
    "windhip_divergence", "windhip_get_lambda", "windhip_set_lambda", "windhip_probe_operator",
    "windhip_time_kernel", "windhip_memory", "windhip_set_alpha", "windhip_apply_operator",
    "windhip_get_rows", "windhip_profile",
    "windhip_net_param_count", "windhip_net_create", "windhip_net_destroy", "windhip_net_last_error",
    "windhip_net_set_params", "windhip_net_get_params", "windhip_net_forward", "windhip_net_loss_grad",
    "windhip_net_adam_step", "windhip_net_time_forward", "windhip_set_net", "windhip_learned_input", "windhip_set_guess_smoothing",
    "windhip_set_multigrid", "windhip_mg_level", "windhip_mg_smooth",
    "windhip_balance_uvw", "windhip_calc_w_real", "windhip_idealized_wgrid",
    "windhip_balance_uvw_noalloc",
    "windhip_probe_lambda_pattern", "windhip_probe_record", "windhip_probe_finalize",
    "windhip_mg_cycle", "windhip_set_mg_fusion", "windhip_set_mg_galerkin", "windhip_mg_stencil",
)

PROFILE_KINDS = {0: "sweep_f64", 1: "sweep_sigma_f64", 2: "update_s_f64", 3: "sweep_omega_f64",
                 4: "update_xrp_f64", 5: "sweep_f32", 6: "sweep_sigma_f32", 7: "update_s_f32",
                 8: "sweep_omega_f32", 9: "update_xrp_f32",
                 10: "init_f64", 11: "ir_resid", 12: "ir_accumulate", 13: "true_resid",
                 14: "fg_setup", 15: "fg_matvec", 16: "fg_gram_schmidt", 17: "line_solve_f32",
                 18: "resid_f32", 19: "mv_sigma_f32", 20: "mv_omega_f32", 21: "mg_transfer_f32",
                 22: "mg_coarse_f32", 23: "mg_setup", 24: "smooth_f32",
                 25: "resid_restrict_f32"}
PROFILE_SLOTS = 26  # WINDHIP_PROFILE_KINDS


class Desc(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("dx", C.c_double),
                ("ims", C.c_int), ("ime", C.c_int), ("kms", C.c_int), ("kme", C.c_int),
                ("jms", C.c_int), ("jme", C.c_int), ("field_dtype", C.c_int), ("field_mem", C.c_int)]


class SolveOpts(C.Structure):
    _fields_ = [("tol_rel", C.c_double), ("tol_abs", C.c_double), ("max_iters", C.c_int),
                ("start", C.c_int), ("precision", C.c_int), ("inner_tol", C.c_double),
                ("inner_max_iters", C.c_int), ("max_outer", C.c_int), ("iters_per_sync", C.c_int),
                ("use_graph", C.c_int), ("retry", C.c_int), ("profile", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("status", C.c_int), ("iters", C.c_int), ("nsweeps", C.c_int), ("outer_iters", C.c_int),
                ("res0", C.c_double), ("res_final", C.c_double), ("bnorm", C.c_double),
                ("true_rel_res", C.c_double), ("ms", C.c_double), ("start_used", C.c_int),
                ("reserved", C.c_int), ("guess_res", C.c_double), ("plain_res", C.c_double)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class WindHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"windhip error {code}: {msg}")
        self.code = code


_lib = None


def _init_torch_runtime_first() -> None:
    """PyTorch's ROCm wheel carries its own HIP/HSA runtime next to the system
    one this library links. In one process torch only finds the GPU if its
    runtime initialises before the system one, so when torch is present it is
    initialised before the library is loaded. Hosts without torch see a
    single runtime and skip this."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except RuntimeError:
        pass


def load():
    """Load libwindhip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hicar_amd/csrc/Makefile)")
    _init_torch_runtime_first()
    lib = C.CDLL(LIB_PATH)
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    pp = C.POINTER(C.c_void_p)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    sig = {
        "windhip_default_opts": (None, [C.POINTER(SolveOpts)]),
        "windhip_create": (i, [C.POINTER(Desc), pp]),
        "windhip_destroy": (i, [vp]),
        "windhip_last_error": (C.c_char_p, []),
        "windhip_set_metrics": (i, [vp] + [vp] * 15),
        "windhip_solve": (i, [vp, vp, vp, C.POINTER(SolveOpts), C.POINTER(SolveInfo)]),
        "windhip_apply_correction": (i, [vp, vp, vp, vp, vp, i]),
        "windhip_calc_iter_winds": (i, [vp, vp, vp, vp, vp, vp, vp, i, C.POINTER(SolveOpts),
                                        C.POINTER(SolveInfo)]),
        "windhip_divergence": (i, [vp, vp, vp, vp, vp, i, i, i, vp]),
        "windhip_get_lambda": (i, [vp, dp]),
        "windhip_set_lambda": (i, [vp, dp]),
        "windhip_probe_operator": (i, [vp, vp, vp, i, dp]),
        "windhip_time_kernel": (i, [vp, i, i, dp]),
        "windhip_memory": (i, [vp, dp, dp]),
        "windhip_set_alpha": (i, [vp, vp]),
        "windhip_apply_operator": (i, [vp, i, dp, dp]),
        "windhip_get_rows": (i, [vp, i, dp]),
        "windhip_profile": (i, [vp, i, dp, dp]),
        "windhip_net_param_count": (C.c_long, []),
        "windhip_net_create": (i, [pp]),
        "windhip_net_destroy": (i, [vp]),
        "windhip_net_last_error": (C.c_char_p, [vp]),
        "windhip_net_set_params": (i, [vp, fp]),
        "windhip_net_get_params": (i, [vp, fp]),
        "windhip_net_forward": (i, [vp, vp, i, i, i, i, vp]),
        "windhip_net_loss_grad": (i, [vp, vp, vp, i, i, i, i, dp, fp]),
        "windhip_net_adam_step": (i, [vp, C.c_float, C.c_float, C.c_float, C.c_float]),
        "windhip_net_time_forward": (i, [vp, i, i, i, i, i, dp]),
        "windhip_set_net": (i, [vp, vp]),
        "windhip_set_guess_smoothing": (i, [vp, i]),
        "windhip_set_multigrid": (i, [vp, i, i, i, i, i, d]),
        "windhip_mg_level": (i, [vp, i, C.POINTER(C.c_int), fp]),
        "windhip_mg_smooth": (i, [vp, i, i, fp, fp, fp]),
        "windhip_balance_uvw": (i, [vp, vp, vp, vp, vp, i, i]),
        "windhip_balance_uvw_noalloc": (i, [vp, vp, vp, vp, vp, i, i]),
        "windhip_probe_lambda_pattern": (i, [vp, i, i, i]),
        "windhip_probe_record": (i, [vp, vp, i, i, i, dp]),
        "windhip_probe_finalize": (i, [vp]),
        "windhip_mg_cycle": (i, [vp, i, i, fp, fp]),
        "windhip_set_mg_fusion": (i, [vp, i, i]),
        "windhip_set_mg_galerkin": (i, [vp, i]),
        "windhip_mg_stencil": (i, [vp, i, fp]),
        "windhip_calc_w_real": (i, [vp, vp, vp, vp, vp]),
        "windhip_idealized_wgrid": (i, [vp, vp, vp, vp, vp]),
        "windhip_learned_input": (i, [vp, vp, C.POINTER(C.c_int), dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != WINDHIP_OK:
        msg = load().windhip_last_error()
        raise WindHipError(rc, msg.decode() if msg else "")


def default_opts(**kw) -> SolveOpts:
    o = SolveOpts()
    load().windhip_default_opts(C.byref(o))
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o
