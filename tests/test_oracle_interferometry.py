"""Pins the gather-construction oracle (oracle/interferometry_oracle.py) with
what the mathematics fixes: the FFT correlation equals the brute-force lag sum
(in float64 to round-off, in float32 to the FFT's rounding) with the
reference's sign convention, central differences are exact on quadratics,
the spectral derivative of a periodic band-limited signal is its analytic
derivative, linear interpolation reproduces linear functions and the sample
values, and the branches relate as reversal and average."""
import numpy as np
import pytest

from oracle import interferometry_oracle as IO


def _noise(nt, nch, seed=0):
    return np.random.default_rng(seed).standard_normal((nt, nch)).astype(np.float32)


@pytest.mark.parametrize("nt,nlag", [(50, 7), (64, 63), (33, 0), (17, 20)])
def test_fft_correlation_equals_lag_sum(nt, nlag):
    x = _noise(nt, 6, seed=nt)
    rec = np.array([0, 2, 3, 5])
    dt = 0.01
    fast, lags = IO.cross_correlation(x, 1, nlag * dt, dt, rec, work_dtype=np.float64)
    slow = IO.direct_correlation(x, 1, nlag, rec)
    np.testing.assert_allclose(fast, slow.astype(np.float32), rtol=1e-6, atol=1e-5)
    f32, _ = IO.cross_correlation(x, 1, nlag * dt, dt, rec)
    err = np.linalg.norm(f32 - slow) / np.linalg.norm(slow)
    assert err < 1e-5
    assert len(lags) == 2 * nlag + 1 and lags[nlag] == 0.0


def test_correlation_sign_convention():
    # receiver = source delayed by d samples: the correlation peaks at lag +d
    nt, d = 200, 9
    s = np.random.default_rng(1).standard_normal(nt).astype(np.float32)
    x = np.zeros((nt, 2), np.float32)
    x[:, 0] = s
    x[d:, 1] = s[:nt - d]
    c, lags = IO.cross_correlation(x, 0, 20 * 0.01, 0.01, np.array([1]))
    assert int(np.argmax(c[:, 0])) - 20 == d


def test_central_difference_exact_on_quadratics():
    t = np.arange(12) * 0.5
    f = (3.0 * t ** 2 - 2.0 * t + 1.0)[:, None].astype(np.float32)
    d = IO.time_derivative(f, 0.5, work_dtype=np.float64)
    np.testing.assert_allclose(d[1:-1, 0], 6.0 * t[1:-1] - 2.0, rtol=1e-6)
    # one-sided first-order edges
    assert d[0, 0] == pytest.approx((f[1, 0] - f[0, 0]) / 0.5, rel=1e-6)
    assert d[-1, 0] == pytest.approx((f[-1, 0] - f[-2, 0]) / 0.5, rel=1e-6)


def test_spectral_derivative_of_periodic_sinusoid():
    n, dt, k = 64, 0.01, 5
    t = np.arange(n) * dt
    w = 2 * np.pi * k / (n * dt)
    f = np.sin(w * t)[:, None]
    d = IO.time_derivative(f, dt, scheme="spectral", work_dtype=np.float64)
    np.testing.assert_allclose(d[:, 0], w * np.cos(w * t), atol=1e-4 * w)


def test_linear_interpolation():
    g = (np.arange(40) * 0.25 - 3.0)[:, None] * np.array([[1.0, -2.0]], np.float32)
    g = g.astype(np.float32)
    out = IO.resample(g, 0.02, 0.0075, 120)
    t = np.arange(120) * 0.0075
    inside = t <= 39 * 0.02
    want = ((t / 0.02) * 0.25 - 3.0)[:, None] * np.array([[1.0, -2.0]])
    np.testing.assert_allclose(out[inside], want[inside], rtol=1e-5, atol=1e-5)
    assert not out[~inside].any()
    same = IO.resample(g, 0.02, 0.02, 40)
    np.testing.assert_array_equal(same, g)


def test_branches():
    lags = np.arange(-5, 6) * 0.1
    p = np.random.default_rng(3).standard_normal((11, 3)).astype(np.float32)
    c, tc = IO.branch(p, lags, "causal")
    a, ta = IO.branch(p, lags, "anticausal")
    s, ts = IO.branch(p, lags, "symmetric")
    np.testing.assert_array_equal(c, p[5:])
    np.testing.assert_array_equal(a, p[5::-1])
    np.testing.assert_array_equal(s, 0.5 * (c + a))
    np.testing.assert_allclose(ta, tc)
    even = np.concatenate([p[5:][::-1], p[6:]])
    np.testing.assert_array_equal(IO.branch(even, lags, "symmetric")[0], IO.branch(even, lags)[0])
