"""Preprocessing of continuous DAS data before correlation, on the GPU.

The reference imports these five functions (run_das_pipeline.py:11-17, used
at :165-168 and :280) from a module whose source is absent from its tree; the
behaviour follows its README (README.md:27-52) and the scipy routines named
there.  Choices the README leaves open are stated per function.  Filters are
designed on the host with scipy.signal (a few coefficients); every sample is
filtered, normalised and transformed in HIP kernels
(das_amd/csrc/dsp.hip) -- the zero-phase filter reproduces
scipy.signal.sosfiltfilt's FP64 arithmetic bit for bit.
"""
from __future__ import annotations

import numpy as np
from scipy import signal

from . import _native as N


def _sos_args(sos, nt):
    """(sos, zi, nsec, padlen) for the zero-phase filter of nt samples."""
    sos = np.ascontiguousarray(sos, dtype=np.float64)
    zi = np.ascontiguousarray(signal.sosfilt_zi(sos), dtype=np.float64)
    nsec = sos.shape[0]
    # scipy.signal.sosfiltfilt's default pad length
    padlen = 3 * (2 * nsec + 1 - min(int((sos[:, 2] == 0).sum()), int((sos[:, 5] == 0).sum())))
    if nt <= padlen:
        raise ValueError(f"The length of the input vector x must be greater than padlen, "
                         f"which is {padlen}.")
    return sos, zi, nsec, padlen


def _sosfiltfilt(x, sos, step=1, sign=False):
    x = N.f32(x)
    if x.ndim == 1:
        return _sosfiltfilt(x[:, None], sos, step, sign)[:, 0]
    nt, nch = x.shape
    sos, zi, nsec, padlen = _sos_args(sos, nt)
    out = np.empty(((nt + step - 1) // step, nch), np.float32)
    N.check(N.lib().tsw_sosfiltfilt(N.ptr(x), nt, nch, N.ptr(sos), nsec, N.ptr(zi), padlen,
                                    int(step), 1 if sign else 0, N.ptr(out)))
    return out


def decimate_time(data: np.ndarray, fs_orig: float, target_fs: float):
    """README.md:27-33: Butterworth low-pass at the new Nyquist frequency
    (order 8, zero-phase) then every q-th sample, q = round(fs_orig /
    target_fs).  Returns (data (nt/q, nch) float32, dt_new, fs_new).  With
    q <= 1 the data are returned unfiltered."""
    q = int(round(fs_orig / target_fs))
    if q <= 1:
        return N.f32(data).copy(), 1.0 / fs_orig, float(fs_orig)
    sos = signal.butter(8, (fs_orig / q) / 2.0, btype="lowpass", fs=fs_orig, output="sos")
    fs_new = fs_orig / q
    return _sosfiltfilt(data, sos, step=q), 1.0 / fs_new, fs_new


def bandpass_filter(data: np.ndarray, fs: float, fmin: float, fmax: float,
                    order: int = 4) -> np.ndarray:
    """README.md:35-38: 4th-order Butterworth band-pass, zero-phase
    (sosfiltfilt) along time; float32 out."""
    sos = signal.butter(order, [fmin, fmax], btype="bandpass", fs=fs, output="sos")
    return _sosfiltfilt(data, sos)


def temporal_normalization(data: np.ndarray, method: str = "1bit") -> np.ndarray:
    """README.md:40-44: 1-bit normalisation, np.sign(data).astype(float32)."""
    if method != "1bit":
        raise ValueError(f"unknown temporal normalization: {method}")
    x = N.f32(data)
    out = np.empty_like(x)
    N.check(N.lib().tsw_sign(N.ptr(x), x.size, N.ptr(out)))
    return out


def spectral_whitening(data: np.ndarray, fs: float, fmin: float, fmax: float,
                       eps: float = 1e-10) -> np.ndarray:
    """README.md:46-52: along time, X(f) / (|X(f)| + eps) for fmin <= f <=
    fmax and 0 outside (the README's "same energy inside the band"), float32
    transforms; eps (absolute) is not given by the README."""
    x = N.f32(data)
    squeeze = x.ndim == 1
    if squeeze:
        x = x[:, None]
    nt, nch = x.shape
    out = np.empty_like(x)
    N.check(N.lib().tsw_spectral_whitening(N.ptr(x), nt, nch, float(fs), float(fmin),
                                           float(fmax), float(eps), N.ptr(out)))
    return out[:, 0] if squeeze else out


def fk_filter(gather: np.ndarray, dt: float, dx: float, vmin: float, vmax: float) -> np.ndarray:
    """F-K fan filter of a (nt, n_receivers) gather (run_das_pipeline.py:280):
    keeps the wavenumbers k != 0 whose apparent velocity |f| / |k| lies in
    [vmin, vmax], both propagation directions, without taper (the README does
    not describe a taper); float32 transforms."""
    g = N.f32(gather)
    nt, nx = g.shape
    out = np.empty_like(g)
    N.check(N.lib().tsw_fk_filter(N.ptr(g), nt, nx, float(dt), float(dx), float(vmin),
                                  float(vmax), N.ptr(out)))
    return out


def preprocess_blocks_device(blocks, fs_orig: float, target_fs: float, fmin: float, fmax: float,
                             outs=None, eps: float = 1e-10):
    """decimate_time, bandpass_filter, temporal_normalization("1bit") and
    spectral_whitening (run_das_pipeline.py:165-168) of several (nt, nch)
    blocks of one shape in one call: each block is copied to the device once
    and its result stays there, and the zero-phase filters run the channels of
    all the blocks side by side.  -> (list of torch float32 CUDA tensors
    (nt_proc, nch), dt_proc, fs_proc); each block bit-identical to the four
    functions in turn.  ``outs`` (CUDA float32 tensors of that shape) are
    reused when given."""
    import torch

    xs = [N.f32(b) for b in blocks]
    if not xs:
        raise ValueError("no blocks")
    if any(x.ndim != 2 for x in xs):
        raise ValueError("blocks must be (nt, n_channels)")
    nt, nch = xs[0].shape
    if any(x.shape != (nt, nch) for x in xs):
        raise ValueError("blocks of one call must have one shape")
    q = int(round(fs_orig / target_fs))
    lp = (None, None, 0, 0)
    if q > 1:
        fs_proc = fs_orig / q
        lp = _sos_args(signal.butter(8, fs_proc / 2.0, btype="lowpass", fs=fs_orig,
                                     output="sos"), nt)
    else:
        q, fs_proc = 1, float(fs_orig)
    nt1 = (nt + q - 1) // q
    bp = _sos_args(signal.butter(4, [fmin, fmax], btype="bandpass", fs=fs_proc, output="sos"),
                   nt1)
    given = list(outs) if outs is not None else []

    def fits(o):
        return (o is not None and tuple(o.shape) == (nt1, nch) and o.dtype == torch.float32
                and o.is_cuda and o.is_contiguous())

    outs = [given[i] if i < len(given) and fits(given[i])
            else torch.empty((nt1, nch), dtype=torch.float32, device="cuda")
            for i in range(len(xs))]
    xp = (N.P * len(xs))(*[N.ptr(x) for x in xs])
    op = (N.P * len(xs))(*[o.data_ptr() for o in outs])
    ptr = lambda a: None if a is None else N.ptr(a)  # noqa: E731
    N.check(N.lib().tsw_preprocess_device_blocks(
        len(xs), xp, nt, nch, ptr(lp[0]), lp[2], ptr(lp[1]), lp[3], q, N.ptr(bp[0]), bp[2],
        N.ptr(bp[1]), bp[3], float(fs_proc), float(fmin), float(fmax), float(eps), op,
        torch.cuda.current_stream(outs[0].device).cuda_stream))
    return outs, 1.0 / fs_proc, fs_proc


def preprocess_block_device(data: np.ndarray, fs_orig: float, target_fs: float, fmin: float,
                            fmax: float, out=None, eps: float = 1e-10):
    """preprocess_blocks_device of one block -> (torch float32 CUDA tensor
    (nt_proc, nch), dt_proc, fs_proc); ``out`` is reused when it fits."""
    x = N.f32(data)
    if x.ndim != 2:
        raise ValueError("data must be (nt, n_channels)")
    outs, dt, fs = preprocess_blocks_device([x], fs_orig, target_fs, fmin, fmax,
                                            outs=None if out is None else [out], eps=eps)
    return outs[0], dt, fs
