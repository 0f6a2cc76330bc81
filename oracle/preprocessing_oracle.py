"""CPU oracle for the preprocessing chain (TEST INFRASTRUCTURE ONLY).

The reference's preprocessing module is absent (bytecode only); this restates
its README (README.md:27-52) with the scipy / numpy routines it names, with
the same choices das_amd.preprocessing documents (decimation filter order 8
at the new Nyquist, whitening eps, untapered F-K fan).  Parity to the
reference itself is therefore unpinned; tests/test_oracle_preprocessing.py
pins this restatement by filter responses and closed forms.
"""
import numpy as np
from scipy import signal


def decimate_time(data, fs_orig, target_fs):
    q = int(round(fs_orig / target_fs))
    x = np.asarray(data, np.float32)
    if q <= 1:
        return x.copy(), 1.0 / fs_orig, float(fs_orig)
    sos = signal.butter(8, (fs_orig / q) / 2.0, btype="lowpass", fs=fs_orig, output="sos")
    y = signal.sosfiltfilt(sos, x, axis=0)[::q]
    return y.astype(np.float32), q / fs_orig, fs_orig / q


def bandpass_filter(data, fs, fmin, fmax, order=4):
    sos = signal.butter(order, [fmin, fmax], btype="bandpass", fs=fs, output="sos")
    return signal.sosfiltfilt(sos, np.asarray(data, np.float32), axis=0).astype(np.float32)


def temporal_normalization(data):
    return np.sign(data).astype(np.float32)


def spectral_whitening(data, fs, fmin, fmax, eps=1e-10, work_dtype=np.float32):
    x = np.asarray(data, work_dtype)
    nt = x.shape[0]
    X = np.fft.rfft(x, axis=0)
    f = np.fft.rfftfreq(nt, d=1.0 / fs)
    band = ((f >= fmin) & (f <= fmax))[:, None] if x.ndim == 2 else (f >= fmin) & (f <= fmax)
    W = np.where(band, X / (np.abs(X) + eps), 0)
    return np.fft.irfft(W, n=nt, axis=0).astype(np.float32)


def fk_filter(gather, dt, dx, vmin, vmax, work_dtype=np.complex64):
    g = np.asarray(gather)
    nt, nx = g.shape
    X = np.fft.fft2(g.astype(work_dtype))
    f = np.abs(np.fft.fftfreq(nt, d=dt))[:, None]
    k = np.abs(np.fft.fftfreq(nx, d=dx))[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        v = np.where(k > 0, f / np.where(k > 0, k, 1.0), np.inf)
    keep = (k > 0) & (v >= vmin) & (v <= vmax)
    return np.real(np.fft.ifft2(np.where(keep, X, 0))).astype(np.float32)
