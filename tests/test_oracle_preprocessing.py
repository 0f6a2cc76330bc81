"""Pins the preprocessing restatement (oracle/preprocessing_oracle.py; the
reference's own module is absent, so parity to it is unpinned) by what the
filters must do: pass in-band sinusoids with unit gain and zero phase,
remove out-of-band ones, decimate a low-frequency sinusoid to its samples,
1-bit is the sign, whitening flattens the in-band spectrum and empties the
rest, the F-K fan keeps a plane wave inside the velocity window and removes
one outside."""
import numpy as np
import pytest

from oracle import preprocessing_oracle as PO


def _sin(f, fs, nt, phase=0.3):
    t = np.arange(nt) / fs
    return np.sin(2 * np.pi * f * t + phase)


def test_bandpass_unit_gain_zero_phase_in_band():
    fs, nt = 50.0, 6000
    x = np.stack([_sin(2.0, fs, nt), _sin(0.1, fs, nt), _sin(15.0, fs, nt)], axis=1)
    y = PO.bandpass_filter(x.astype(np.float32), fs, 0.8, 4.5)
    mid = slice(1000, 5000)
    np.testing.assert_allclose(y[mid, 0], x[mid, 0], atol=2e-3)   # unit gain, zero phase
    assert np.abs(y[mid, 1]).max() < 2e-2 and np.abs(y[mid, 2]).max() < 1e-3


def test_decimation_keeps_low_band():
    fs, nt = 1000.0, 20000
    x = np.stack([_sin(3.0, fs, nt), _sin(180.0, fs, nt)], axis=1).astype(np.float32)
    y, dt, fs2 = PO.decimate_time(x, fs, 50.0)
    assert fs2 == 50.0 and dt == pytest.approx(0.02) and y.shape == (1000, 2)
    mid = slice(100, 900)
    np.testing.assert_allclose(y[mid, 0], x[::20][mid, 0], atol=2e-3)
    assert np.abs(y[mid, 1]).max() < 1e-3


def test_one_bit_and_whitening():
    x = np.array([[-2.0, 0.0, 3.5], [np.inf, -0.0, -1e-30]], np.float32)
    np.testing.assert_array_equal(PO.temporal_normalization(x),
                                  np.array([[-1, 0, 1], [1, 0, -1]], np.float32))
    rng = np.random.default_rng(0)
    fs, nt = 50.0, 4096
    x = (rng.standard_normal((nt, 3)) * np.array([1.0, 30.0, 0.01])).astype(np.float32)
    w = PO.spectral_whitening(x, fs, 0.8, 4.5, work_dtype=np.float64)
    W = np.abs(np.fft.rfft(w.astype(np.float64), axis=0))
    f = np.fft.rfftfreq(nt, 1 / fs)
    band = (f >= 0.8) & (f <= 4.5)
    np.testing.assert_allclose(W[band], 1.0, rtol=1e-6)
    assert W[~band].max() < 1e-6


def test_fk_fan_keeps_window_velocities():
    nt, nx, dt, dx = 512, 128, 0.02, 8.0
    t = np.arange(nt)[:, None] * dt
    x = np.arange(nx)[None, :] * dx
    # plane waves on exact FFT bins: f = 40/(nt dt), k = 4/(nx dx) -> v = 1250 m/s;
    # f = 40/(nt dt), k = 20/(nx dx) -> v = 250 m/s
    fa, ka = 40 / (nt * dt), 4 / (nx * dx)
    fb, kb = 40 / (nt * dt), 20 / (nx * dx)
    inside = np.cos(2 * np.pi * (fa * t - ka * x))
    outside = np.cos(2 * np.pi * (fb * t - kb * x))
    y = PO.fk_filter((inside + outside).astype(np.float32), dt, dx, 800.0, 3500.0,
                     work_dtype=np.complex128)
    np.testing.assert_allclose(y, inside, atol=1e-6)
