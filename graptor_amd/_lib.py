"""ctypes binding of libgraptor_amd.so (the C ABI declared in include/graptor_amd.h).

The library must have been built (``python -m graptor_amd.build``); loading
fails loudly otherwise.  There is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libgraptor_amd.so")

GR_OK = 0
GR_ERR_INVALID = 1
GR_ERR_OOM = 2
GR_ERR_HIP = 3
GR_ERR_UNSUPPORTED = 4

COUPLING = {"uncorrelated": 0, "frobenius": 1, "nuclear": 2}
DISCREPANCY = {"KL": 0, "L2": 1}
REGULARISATION = {"TV": 0, "TGV": 1}
NORMEST_MODE = {"reference": 0, "power": 1}
HALO_PHASE = {"dual": 0, "primal": 1}

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = (
    "gr_last_error", "gr_abi_version", "gr_radon_struct", "gr_geometry_create",
    "gr_geometry_destroy", "gr_geometry_set_tuning", "gr_geometry_get_tuning",
    "gr_geometry_set_fusion",
    "gr_radon", "gr_radon_ad", "gr_normest", "gr_make_feasible", "gr_tgv_iterate",
    "gr_tgv_energy", "gr_tgv_profile", "gr_plan_create", "gr_plan_destroy",
    "gr_plan_iterate", "gr_plan_iterate_graph", "gr_plan_profile", "gr_plan_retune",
    "gr_cnn_conv_fwd", "gr_cnn_conv_wgrad", "gr_cnn_bias_grad", "gr_cnn_weight_flip",
    "gr_cnn_relu_bwd", "gr_cnn_pool", "gr_cnn_upsample", "gr_cnn_pack_bf16", "gr_cnn_mse_grad",
    "gr_cnn_adam", "gr_tgv_dual_step", "gr_tgv_primal_step", "gr_csr_apply", "gr_prox",
    "gr_sumsq_f64", "gr_tgv_pack_halos", "gr_tgv_energy_noalloc",
)
ENERGY_NOALLOC_MAX_CHANNELS = 32


class GraptorError(RuntimeError):
    """A failing call into libgraptor_amd."""


class GraptorMemoryError(GraptorError, MemoryError):
    """Device allocation failure; callers that retry with smaller slabs catch
    MemoryError exactly as the reference's caller catches cl.MemoryError."""


class TgvParams(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int),
        ("nz", ctypes.c_int), ("nch", ctypes.c_int), ("coupling", ctypes.c_int),
        ("discrepancy", ctypes.c_int), ("regularisation", ctypes.c_int),
        ("stepsize", ctypes.c_float), ("sigma", ctypes.c_float), ("tau", ctypes.c_float),
        ("sigma_p", ctypes.c_float), ("norming", ctypes.c_float),
        ("alpha_inv0", ctypes.c_float), ("alpha_inv1", ctypes.c_float),
    ]


class TgvState(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in
                ("u", "ubar", "p", "pbar", "v", "w", "lam", "ktl", "ku", "f", "mfac",
                 "u2", "v2", "w2", "halo_ubar_hi", "halo_pbar_lo", "halo_v_lo", "halo_w_hi")]


_lib = None
_lock = threading.Lock()


def load(path: str | None = None):
    """Load the library once and declare its signatures (GRAPTOR_AMD_LIB, if
    set, names another build of it)."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or os.environ.get("GRAPTOR_AMD_LIB") or LIB_PATH
        if not os.path.exists(p):
            raise ImportError(
                f"libgraptor_amd.so not found at {p}; build it with "
                "`python -m graptor_amd.build`")
        lib = ctypes.CDLL(p)
        V, I, D, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
        S = ctypes.c_size_t
        PD = ctypes.POINTER(ctypes.c_double)
        PI = ctypes.POINTER(ctypes.c_int)
        sig = {
            "gr_last_error": ([], ctypes.c_char_p),
            "gr_abi_version": ([], I),
            "gr_radon_struct": ([V, I, I, I, I, D, D, V], I),
            "gr_geometry_create": ([V, I, I, I, I, ctypes.POINTER(V)], I),
            "gr_geometry_destroy": ([V], I),
            "gr_geometry_set_tuning": ([V, I, I], I),
            "gr_geometry_get_tuning": ([V, PI, PI], I),
            "gr_geometry_set_fusion": ([V, I], I),
            "gr_radon": ([V, V, V, I, V], I),
            "gr_radon_ad": ([V, V, V, I, V], I),
            "gr_normest": ([V, V, I, I, D, PD, V], I),
            "gr_make_feasible": ([V, V, I, V], I),
            "gr_tgv_iterate": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState),
                                I, I, V], I),
            "gr_tgv_energy": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState),
                               V, D, D, D, V, V], I),
            "gr_tgv_energy_noalloc": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState),
                                       V, D, D, D, V, V], I),
            "gr_tgv_profile": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState),
                                I, V, V], I),
            "gr_plan_create": ([I, V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState),
                                ctypes.POINTER(V)], I),
            "gr_plan_destroy": ([V], I),
            "gr_plan_iterate": ([V, I, I, V], I),
            "gr_plan_iterate_graph": ([V, I, I, V], I),
            "gr_plan_profile": ([V, I, V, V], I),
            "gr_plan_retune": ([V, V], I),
            "gr_cnn_conv_fwd": ([V, I, I, V, V, V, I, I, I, I, I, I, I, I, I, I, V], I),
            "gr_cnn_conv_wgrad": ([V, I, I, V, I, I, V, I, I, I, I, I, I, I, V], I),
            "gr_cnn_bias_grad": ([V, I, I, S, I, V, V], I),
            "gr_cnn_weight_flip": ([V, V, I, I, I, V], I),
            "gr_cnn_relu_bwd": ([V, I, I, V, I, I, S, I, V], I),
            "gr_cnn_pool": ([V, I, I, V, I, I, I, I, I, I, I, I, V], I),
            "gr_cnn_upsample": ([V, I, I, V, I, I, I, I, I, I, I, V], I),
            "gr_cnn_pack_bf16": ([V, I, V, I, I, S, F, V], I),
            "gr_cnn_mse_grad": ([V, I, V, I, S, V, I, F, V, V], I),
            "gr_cnn_adam": ([V, V, V, V, V, S, F, F, F, F, I, F, V], I),
            "gr_tgv_dual_step": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState), V], I),
            "gr_tgv_primal_step": ([V, ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState), V],
                                   I),
            "gr_csr_apply": ([V, V, V, I, I, I, I, I, I, V, V, V], I),
            "gr_prox": ([I, F, V, I, S, V, V, V, V], I),
            "gr_sumsq_f64": ([V, S, PD, V], I),
            "gr_tgv_pack_halos": ([ctypes.POINTER(TgvParams), ctypes.POINTER(TgvState), I, V, V,
                                   V], I),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        if path is None:
            _lib = lib
        return lib


def check(rc: int, what: str = "") -> None:
    """Raise the exception that matches a status code."""
    if rc == GR_OK:
        return
    msg = load().gr_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == GR_ERR_OOM:
        raise GraptorMemoryError(text)
    if rc == GR_ERR_INVALID:
        raise ValueError(text)
    if rc == GR_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    raise GraptorError(text)
