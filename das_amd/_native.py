"""ctypes binding of libtswrtm_hip.so (include/tswrtm.h).

The library is the only compute path: if it is missing, or no gfx950 device
is present, calls raise instead of falling back to anything else.  This plays
the role of the module-level ``ctypes.CDLL`` load in tsw_rtm/wave_solver.py:14-71.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libtswrtm_hip.so")

P = ctypes.c_void_p
I = ctypes.c_int
F = ctypes.c_float
D = ctypes.c_double
LL = ctypes.c_longlong

# name -> (restype, argtypes); mirrors include/tswrtm.h
SIGNATURES = {
    "acoustic_2d_step": (None, [P] * 5 + [I, I, F, F, F, I]),
    "elastic_2d_step": (None, [P] * 10 + [I, I, F, F, F, I]),
    "correlate_imaging_condition": (None, [P, P, P, I, I, F]),
    "build_cerjan_sponge": (None, [P, I, I, I, F, I]),
    "tsw_last_error": (ctypes.c_char_p, []),
    "tsw_version": (I, []),
    "tsw_init": (I, [I]),
    "tsw_debug_fail_next": (None, [I]),
    "tsw_mem_info": (I, [ctypes.POINTER(LL), ctypes.POINTER(LL)]),
    "tsw_device_name": (I, [ctypes.c_char_p, I]),
    "tsw_grid_create": (I, [I, I, D, D, D, I, D, I, ctypes.POINTER(P)]),
    "tsw_grid_destroy": (I, [P]),
    "tsw_grid_sponge": (I, [P, P]),
    "tsw_grid_set_sponge": (I, [P, P]),
    "tsw_forward_acoustic": (I, [P, P, P, I, I, I, P, I, I, I, P, P]),
    "tsw_backward_correlate": (I, [P, P, P, I, P, I, I, P, LL, P, I]),
    "tsw_migrate_shot": (I, [P, P, P, P, I, I, I, P, I, P, I, I, I, P, P]),
    "tsw_migrate_shots": (I, [P, P, P, I, P, P, P, P, P, P, P, I, I, I, I, P, P, P]),
    "tsw_laplacian": (I, [P, P, I, I, D, D]),
    "tsw_forward_elastic": (I, [P] * 6 + [I, I, I, P, I, I, I, P, P, P]),
    "tsw_backward_correlate_elastic": (I, [P] * 6 + [I, P, I, I, P, LL, P, I]),
    "tsw_migrate_shot_elastic": (I, [P, P, P, P, I, I, P, I, P, I, I, I, P, P]),
    "tsw_batch_create": (I, [I, P, D, D, D, I, D, I, I, I, I, I, ctypes.POINTER(P)]),
    "tsw_kernel_configs": (I, []),
    "tsw_kernel_config": (I, [I] + [ctypes.POINTER(I)] * 4),
    "tsw_kernel_name": (I, [I, ctypes.c_char_p, I]),
    "tsw_batch_run": (I, [P, I]),
    "tsw_batch_time_kernels": (I, [P, I, ctypes.POINTER(D), ctypes.POINTER(I)]),
    "tsw_batch_shot_result": (I, [P, I, P, P]),
    "tsw_batch_stack": (I, [P, I, I, P, P]),
    "tsw_batch_stack_device": (I, [P, I, I, P, P, P]),
    "tsw_batch_info": (I, [P, ctypes.POINTER(LL), ctypes.POINTER(LL), ctypes.POINTER(I),
                           ctypes.POINTER(I), ctypes.POINTER(LL)]),
    "tsw_batch_destroy": (I, [P]),
    "tsw_server_create": (I, [D, D, D, I, D, I, I, LL, I, I, I, ctypes.POINTER(P)]),
    "tsw_server_submit": (I, [P, P, ctypes.POINTER(LL)]),
    "tsw_server_poll": (I, [P, LL, ctypes.POINTER(I)] + [ctypes.POINTER(D)] * 3),
    "tsw_server_wait": (I, [P, LL, D]),
    "tsw_server_result": (I, [P, LL, P, P]),
    "tsw_server_cancel": (I, [P, LL]),
    "tsw_server_stats": (I, [P] + [ctypes.POINTER(LL)] * 7 + [ctypes.POINTER(D)]),
    "tsw_server_destroy": (I, [P]),
    "tsw_server_set_models": (I, [P, I, I, P, P, P, P]),
    "tsw_fwi_create": (I, [I, I, D, D, D, I, D, I, I, I, P, I, ctypes.POINTER(P)]),
    "tsw_fwi_forward": (I, [P, P, P, ctypes.POINTER(D)]),
    "tsw_fwi_backward": (I, [P, P, P, P]),
    "tsw_fwi_set_blocking": (I, [P, I]),
    "tsw_fwi_set_adjoint_blocking": (I, [P, I]),
    "tsw_fwi_set_adjoint_steps": (I, [P, I]),
    "tsw_fwi_predicted": (I, [P, I, P]),
    "tsw_fwi_set_residual": (I, [P, I, P]),
    "tsw_fwi_probe": (I, [P, I, P]),
    "tsw_fwi_destroy": (I, [P]),
    "tsw_xcorr": (I, [P, I, I, P, I, P, I, I, I, P]),
    "tsw_flood_columns": (I, [P, I, I, I, P, I, D, P]),
    "tsw_sosfiltfilt": (I, [P, I, I, P, I, P, I, I, I, P]),
    "tsw_sign": (I, [P, LL, P]),
    "tsw_spectral_whitening": (I, [P, I, I, D, D, D, D, P]),
    "tsw_preprocess_device": (I, [P, I, I, P, I, P, I, I, P, I, P, I, D, D, D, D, P, P]),
    "tsw_preprocess_device_blocks": (I, [I, P, I, I, P, I, P, I, I, P, I, P, I, D, D, D, D, P, P]),
    "tsw_fk_filter": (I, [P, I, I, D, D, D, D, P]),
    "tsw_save_matrix_csv": (I, [ctypes.c_char_p, P, I, I, I, P, P, ctypes.c_char_p,
                                ctypes.c_char_p]),
    "tsw_xcorr_device": (I, [P, I, I, P, I, P, I, I, I, P, P]),
    "tsw_xcorr_device_noalloc": (I, [P, I, I, P, I, P, I, I, I, P, P]),
    "tsw_xcorr_noalloc_scratch_bytes": (LL, [I, I, I, I]),
    "tsw_xcorr_noalloc_scratch_max": (LL, []),
    "tsw_time_derivative": (I, [P, I, I, D, I, P]),
    "tsw_green_branch": (I, [P, I, I, I, I, P, ctypes.POINTER(I)]),
    "tsw_resample_linear": (I, [P, I, I, D, D, I, P]),
}

RUN_BATCHED = 0
RUN_PLAIN = 1


class ShotDesc(ctypes.Structure):
    """tsw_shot_desc of include/tswrtm.h."""
    _fields_ = [("nx", I), ("nz", I), ("nt", I), ("vel_fwd", P), ("vel_bwd", P), ("src", P),
                ("six", I), ("siz", I), ("data", P), ("rix", P), ("nrec", I), ("riz", I),
                ("x0", I), ("z0", I), ("kind", I), ("ela_fwd", P * 4), ("ela_bwd", P * 4)]


class FwiShot(ctypes.Structure):
    """tsw_fwi_shot of include/tswrtm.h."""
    _fields_ = [("nt", I), ("src", P), ("six", I), ("siz", I), ("rix", P), ("nrec", I),
                ("riz", I), ("d_obs", P)]


SHOT_ACOUSTIC = 0
SHOT_ELASTIC = 1


class TswError(RuntimeError):
    """A coarse entry point returned a negative status."""

    def __init__(self, status, message):
        super().__init__(f"tsw status {status}: {message}")
        self.status = status


class WavefieldDiverged(TswError, FloatingPointError):
    """Status -6: a finished shot's image or illumination holds NaN or Inf,
    which is what an unstable time step (dt above the CFL limit) produces."""


_lib = None
_lock = threading.Lock()


def lib():
    """The loaded engine library (raises if it is not built)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise FileNotFoundError(
                        f"HIP engine not built: {LIB_PATH} (run `python -m das_amd.build`)")
                # torch first: its HIP runtime is then the one the engine's
                # hip* symbols bind to, so the engine and the tensors whose
                # device pointers and streams it takes share one runtime (an
                # engine loaded first binds the system runtime, and torch's own
                # then finds no device)
                try:
                    import torch  # noqa: F401
                except ImportError:
                    pass
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in SIGNATURES.items():
                    fn = getattr(L, name)
                    fn.restype = res
                    fn.argtypes = args
                _lib = L
    return _lib


def last_error() -> str:
    return lib().tsw_last_error().decode()


def check(status):
    """Raise the Python exception matching a negative engine status."""
    if status == 0:
        return
    msg = last_error()
    if status == -1:
        raise ValueError(msg)
    if status == -3:
        raise MemoryError(msg)
    if status == -6:
        raise WavefieldDiverged(status, msg)
    raise TswError(status, msg)


def ptr(a):
    return None if a is None else a.ctypes.data


def f32(a):
    """C-contiguous float32 view or copy (the caller keeps it alive)."""
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
