"""CPU oracle for virtual-shot gather construction.

TEST INFRASTRUCTURE ONLY: the checker that das_amd.interferometry is compared
against; the product never imports it.

Restates, with numpy/scipy, the reference's algorithm:
  - compute_cross_correlation   tsw_rtm/interferometry.py:26-79 (zero-padded
    rfft of length 2^ceil(log2(2nt-1)), conj(src) * rec, irfft, fftshift,
    window of 2 n_lags + 1 around the centre, float32)
  - time_derivative_greens      tsw_rtm/interferometry.py:82-104
  - build_virtual_shot_gather   tsw_rtm/interferometry.py:112-163 (branches)
  - resample_gather_to_sim_dt   run_das_pipeline.py:48-55
Pinned in tests/test_oracle_interferometry.py against brute-force sums and
closed forms.
"""
import numpy as np
from scipy.interpolate import interp1d


def cross_correlation(data, source_channel_idx, max_lag_seconds, dt, receiver_indices=None,
                      work_dtype=None):
    """interferometry.py:26-79.  work_dtype=np.float64 runs the FFTs in double
    (numpy keeps float32 input in single precision)."""
    x = np.asarray(data)
    if work_dtype is not None:
        x = x.astype(work_dtype)
    nt, nch = x.shape
    rec = np.arange(nch) if receiver_indices is None else np.asarray(receiver_indices)
    nfft = int(2 ** np.ceil(np.log2(2 * nt - 1)))
    nlag = int(round(max_lag_seconds / dt))
    spec = np.conjugate(np.fft.rfft(x[:, source_channel_idx], n=nfft))[:, None] * \
        np.fft.rfft(x[:, rec], n=nfft, axis=0)
    full = np.fft.fftshift(np.fft.irfft(spec, n=nfft, axis=0), axes=0)
    c = nfft // 2
    return full[c - nlag:c + nlag + 1, :].astype(np.float32), np.arange(-nlag, nlag + 1) * dt


def direct_correlation(data, source_channel_idx, nlag, receiver_indices):
    """Brute force in float64: c[j, r] = sum_t s[t] * rec_r[t + j - nlag]."""
    x = np.asarray(data, np.float64)
    nt = x.shape[0]
    s = x[:, source_channel_idx]
    out = np.zeros((2 * nlag + 1, len(receiver_indices)))
    for j in range(2 * nlag + 1):
        lag = j - nlag
        t0, t1 = max(0, -lag), min(nt, nt - lag)
        if t1 > t0:
            out[j] = s[t0:t1] @ x[t0 + lag:t1 + lag][:, receiver_indices]
    return out


def time_derivative(ccfs, dt, scheme="central", work_dtype=None):
    """interferometry.py:82-104."""
    c = np.asarray(ccfs)
    if work_dtype is not None:
        c = c.astype(work_dtype)
    if scheme == "central":
        d = np.gradient(c, dt, axis=0)
    else:
        nt = c.shape[0]
        nfft = int(2 ** np.ceil(np.log2(nt)))
        w = 2.0 * np.pi * np.fft.rfftfreq(nfft, d=dt)
        d = np.fft.irfft(1j * w[:, None] * np.fft.rfft(c, n=nfft, axis=0), n=nfft, axis=0)[:nt]
    return d.astype(np.float32)


def branch(processed, lags, which="causal"):
    """interferometry.py:140-155 -> (gather, time_axis)."""
    z = int(np.argmin(np.abs(lags)))
    if which == "causal":
        return processed[z:], lags[z:]
    if which == "anticausal":
        return np.flip(processed[:z + 1], axis=0), np.abs(lags[:z + 1][::-1])
    pos, neg = processed[z:], np.flip(processed[:z + 1], axis=0)
    m = min(len(pos), len(neg))
    return 0.5 * (pos[:m] + neg[:m]), lags[z:z + m]


def resample(gather, dt_orig, dt_sim, target_nt):
    """run_das_pipeline.py:48-55."""
    t_o = np.arange(gather.shape[0]) * dt_orig
    t_s = np.arange(target_nt) * dt_sim
    f = interp1d(t_o, gather, axis=0, kind="linear", bounds_error=False, fill_value=0.0)
    return f(t_s).astype(np.float32)
