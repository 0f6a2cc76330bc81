"""Virtual-shot gathers on the HIP engine: the step before back-injection.

Same names, signatures and return types as tsw_rtm/interferometry.py (and
``resample_gather_to_sim_dt`` of run_das_pipeline.py:48-55); every numeric
step runs in HIP kernels (das_amd/csrc/vgather.hip):

- ``compute_cross_correlation`` (interferometry.py:26-79): the lag window of
  the correlation of the source trace with each receiver trace, summed
  directly in FP64 from exact products (the reference goes through a
  zero-padded FP32 FFT, so the two agree to the FFT's rounding).
- ``time_derivative_greens`` (:82-104): np.gradient's central / one-sided
  differences in FP32, or the spectral derivative.
- ``build_virtual_shot_gather`` (:112-163): derivative, then the causal,
  anticausal or symmetric branch.
- ``resample_gather_to_sim_dt``: scipy interp1d(kind="linear",
  fill_value=0) arithmetic, FP32 differences and FP64 interpolation.

``cross_correlations`` correlates many source channels in one launch and can
stack over files in place, as run_das_pipeline.py:202-213 does per source.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _native as N

_BRANCH = {"causal": 0, "anticausal": 1, "symmetric": 2}


@dataclass
class VirtualShotGather:
    source_idx: int
    source_channel: int
    source_coord_x: float
    receiver_channels: np.ndarray
    receiver_coords_x: np.ndarray
    offsets: np.ndarray
    time_axis: np.ndarray
    shot_gather: np.ndarray  # shape: (nt, n_receivers)
    dt: float


def _n_lags(max_lag_seconds: float, dt: float) -> int:
    return int(round(max_lag_seconds / dt))


def _xcorr_args(nch, source_channels, receiver_indices, max_lag_seconds, dt):
    """Channel lists (negative indices from the end, checked), lag count and
    output shape of the correlations."""
    rec = np.arange(nch) if receiver_indices is None else np.asarray(receiver_indices)
    rec = np.ascontiguousarray(np.where(rec < 0, rec + nch, rec), dtype=np.int32)
    src = np.asarray(source_channels).reshape(-1)
    src = np.ascontiguousarray(np.where(src < 0, src + nch, src), dtype=np.int32)
    for a, what in ((src, "source_channel_idx"), (rec, "receiver_indices")):
        if a.size and (a.min() < 0 or a.max() >= nch):
            raise IndexError(f"{what} out of range for {nch} channels")
    nlag = _n_lags(max_lag_seconds, dt)
    return src, rec, nlag, (len(src), 2 * nlag + 1, len(rec))


def cross_correlations(data: np.ndarray, source_channels: Sequence[int], max_lag_seconds: float,
                       dt: float, receiver_indices: Optional[np.ndarray] = None,
                       out: Optional[np.ndarray] = None) -> Tuple[np.ndarray, np.ndarray]:
    """Correlations of every source channel with the receiver channels:
    (n_src, 2 n_lags + 1, n_rec) float32 and the lag axis.  With ``out``
    (same shape, float32) the correlations are added into it in FP32 (stacking
    over files)."""
    d = N.f32(data)
    if d.ndim != 2:
        raise ValueError("data must be (nt, n_channels)")
    nt, nch = d.shape
    src, rec, nlag, shape = _xcorr_args(nch, source_channels, receiver_indices, max_lag_seconds,
                                        dt)
    acc = 0
    if out is None:
        out = np.empty(shape, np.float32)
    else:
        if out.shape != shape or out.dtype != np.float32 or not out.flags.c_contiguous:
            raise ValueError(f"out must be a C-contiguous float32 array of shape {shape}")
        acc = 1
    N.check(N.lib().tsw_xcorr(N.ptr(d), nt, nch, N.ptr(src), len(src), N.ptr(rec), len(rec), nlag,
                              acc, N.ptr(out)))
    lags = np.arange(-nlag, nlag + 1) * dt
    return out, lags


def cross_correlations_device(data, source_channels: Sequence[int], max_lag_seconds: float,
                              dt: float, receiver_indices: Optional[np.ndarray] = None,
                              out=None):
    """cross_correlations of a (nt, nch) float32 CUDA tensor into a CUDA
    tensor; with ``out`` the correlations are added into it in FP32 (stacking
    over files on the device) -> (out, lags)."""
    import torch

    if data.ndim != 2 or data.dtype != torch.float32 or not data.is_cuda \
            or not data.is_contiguous():
        raise ValueError("data must be a contiguous (nt, n_channels) float32 CUDA tensor")
    nt, nch = data.shape
    src, rec, nlag, shape = _xcorr_args(nch, source_channels, receiver_indices, max_lag_seconds,
                                        dt)
    acc = 0
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=data.device)
    else:
        if tuple(out.shape) != shape or out.dtype != torch.float32 or not out.is_cuda \
                or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float32 CUDA tensor of shape {shape}")
        acc = 1
    N.check(N.lib().tsw_xcorr_device(data.data_ptr(), nt, nch, N.ptr(src), len(src), N.ptr(rec),
                                     len(rec), nlag, acc, out.data_ptr(),
                                     torch.cuda.current_stream(data.device).cuda_stream))
    return out, np.arange(-nlag, nlag + 1) * dt


def compute_cross_correlation(data: np.ndarray, source_channel_idx: int, max_lag_seconds: float,
                              dt: float, receiver_indices: Optional[np.ndarray] = None
                              ) -> Tuple[np.ndarray, np.ndarray]:
    """interferometry.py:26-79 -> (ccfs (2 n_lags + 1, n_rec) float32, lags)."""
    c, lags = cross_correlations(data, [source_channel_idx], max_lag_seconds, dt,
                                 receiver_indices)
    return c[0], lags


def time_derivative_greens(ccfs: np.ndarray, dt: float, scheme: str = "central") -> np.ndarray:
    """interferometry.py:82-104: d(ccf)/dt along axis 0, float32."""
    c = N.f32(ccfs)
    squeeze = c.ndim == 1
    if squeeze:
        c = c[:, None]
    if c.ndim != 2:
        raise ValueError("ccfs must be (n_lags, n_receivers)")
    nl, ncol = c.shape
    out = np.empty_like(c)
    N.check(N.lib().tsw_time_derivative(N.ptr(c), nl, ncol, float(dt),
                                        0 if scheme == "central" else 1, N.ptr(out)))
    return out[:, 0] if squeeze else out


def full_green_function(ccfs: np.ndarray, lags: np.ndarray) -> np.ndarray:
    """Green's function over the full lag axis (interferometry.py:107-109)."""
    return time_derivative_greens(ccfs, float(lags[1] - lags[0]))


def build_virtual_shot_gather(ccfs: np.ndarray, lags: np.ndarray, source_idx: int,
                              source_channel: int, channel_coords_x: np.ndarray,
                              receiver_indices: np.ndarray, branch: str = "causal",
                              take_derivative: bool = True) -> VirtualShotGather:
    """interferometry.py:112-163."""
    dt = float(lags[1] - lags[0])
    zero_idx = int(np.argmin(np.abs(lags)))
    if branch not in _BRANCH:
        raise ValueError(f"Unknown branch option: {branch}")
    processed = time_derivative_greens(ccfs, dt) if take_derivative else N.f32(ccfs)
    nl, ncol = processed.shape
    b = _BRANCH[branch]
    m = nl - zero_idx if b == 0 else (zero_idx + 1 if b == 1 else min(nl - zero_idx, zero_idx + 1))
    gather = np.empty((m, ncol), np.float32)
    nout = N.I(m)
    N.check(N.lib().tsw_green_branch(N.ptr(processed), nl, ncol, zero_idx, b, N.ptr(gather),
                                     ctypes.byref(nout)))
    if b == 0:
        time_axis = lags[zero_idx:]
    elif b == 1:
        time_axis = np.abs(lags[:zero_idx + 1][::-1])
    else:
        time_axis = lags[zero_idx:zero_idx + m]
    src_x = float(channel_coords_x[source_channel])
    rec_x = channel_coords_x[receiver_indices]
    return VirtualShotGather(source_idx=source_idx, source_channel=source_channel,
                             source_coord_x=src_x, receiver_channels=receiver_indices,
                             receiver_coords_x=rec_x, offsets=np.abs(rec_x - src_x),
                             time_axis=time_axis, shot_gather=gather, dt=dt)


def resample_gather_to_sim_dt(gather: np.ndarray, dt_orig: float, dt_sim: float,
                              target_nt: int) -> np.ndarray:
    """run_das_pipeline.py:48-55: linear interpolation of the time axis to the
    simulation step, 0 beyond the gather's time range."""
    g = N.f32(gather)
    if g.ndim != 2:
        raise ValueError("gather must be (nt, n_receivers)")
    out = np.empty((int(target_nt), g.shape[1]), np.float32)
    N.check(N.lib().tsw_resample_linear(N.ptr(g), g.shape[0], g.shape[1], float(dt_orig),
                                        float(dt_sim), int(target_nt), N.ptr(out)))
    return out
