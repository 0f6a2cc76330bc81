"""Host-side functions of the drop-in wave_solver module (no GPU needed):
ricker_wavelet (tsw_rtm/wave_solver.py:74-81) and
WaveSolver2D.check_stability (tsw_rtm/wave_solver.py:119-126)."""
import math

import numpy as np
import pytest

from das_amd.wave_solver import WaveSolver2D, ricker_wavelet
from oracle import oracle as ora


@pytest.mark.parametrize("f0,dt,T", [(2.5, 5e-4, 2.0), (25.0, 5e-4, 0.05), (30.0, 4e-4, 0.3),
                                     (10.0, 1e-3, 0.0125)])
def test_ricker_equals_oracle_and_closed_form(f0, dt, T):
    w, t = ricker_wavelet(f0, dt, T)
    wo, to = ora.ricker(f0, dt, T)
    assert w.dtype == np.float32 and t.dtype == np.float64
    np.testing.assert_array_equal(w, wo)
    np.testing.assert_array_equal(t, to)
    # nt = round(T / dt) (Python's round: half to even), t = n dt
    assert len(w) == int(round(T / dt))
    np.testing.assert_array_equal(t, np.arange(len(w)) * dt)
    # the closed form evaluated sample by sample with the math module
    ref = np.array([(1 - 2 * (math.pi * f0 * (ti - 1 / f0)) ** 2)
                    * math.exp(-(math.pi * f0 * (ti - 1 / f0)) ** 2) for ti in t])
    np.testing.assert_allclose(w, ref, rtol=2e-7, atol=1e-37)


def test_ricker_shape_properties():
    f0, dt = 5.0, 1e-4
    w, t = ricker_wavelet(f0, dt, 1.0)
    i = int(np.argmax(w))
    assert abs(t[i] - 1.0 / f0) <= dt / 2 and w[i] == np.float32(1.0)
    # zero crossings at tau = +-1 / (pi f0 sqrt 2), minima -2 exp(-3/2) at tau = +-sqrt(3/2)/(pi f0)
    tz = 1.0 / (math.pi * f0 * math.sqrt(2.0))
    for s in (-1, 1):
        j = int(round((1.0 / f0 + s * tz) / dt))
        assert abs(w[j]) < 2e-3
    assert abs(float(w.min()) + 2 * math.exp(-1.5)) < 1e-5
    # symmetric about the peak
    np.testing.assert_allclose(w[i - 500:i + 501], w[i - 500:i + 501][::-1], atol=1e-6)


def _solver(dx, dz, dt):
    s = WaveSolver2D.__new__(WaveSolver2D)  # the check reads only dx, dz, dt
    s.dx, s.dz, s.dt = float(dx), float(dz), float(dt)
    return s


def test_check_stability_thresholds_and_message():
    s = _solver(5.0, 5.0, 5e-4)
    ok, msg = s.check_stability(2575.0, 1500.0, 4.5)
    cfl = 2575.0 * 5e-4 * math.sqrt(2.0 / 25.0)
    ppw = (1500.0 / 4.5) / 5.0
    assert ok
    assert msg == (f"CFL number: {cfl:.3f} (target < 0.6). "
                   f"Grid pts per min wavelength: {ppw:.1f} (target >= 5).")
    # the tested thresholds are 0.65 and 4.0, not the 0.6 and 5 the message names
    vmax_065 = 0.65 / (5e-4 * math.sqrt(2.0 / 25.0))
    assert s.check_stability(vmax_065 * 0.999, 1500.0, 4.5)[0]
    assert not s.check_stability(vmax_065 * 1.001, 1500.0, 4.5)[0]
    assert s.check_stability(2575.0, 4.0 * 5.0 * 4.5, 4.5)[0]          # ppw == 4.0 passes
    assert not s.check_stability(2575.0, 3.999 * 5.0 * 4.5, 4.5)[0]
    # CFL 0.62 (above the message's 0.6) and 4.5 points (below its 5) still pass
    ok, msg = s.check_stability(0.62 / (5e-4 * math.sqrt(2.0 / 25.0)), 4.5 * 5.0 * 10.0, 10.0)
    assert ok and "CFL number: 0.620" in msg and "wavelength: 4.5 (" in msg
    # unequal spacing: max(dx, dz) sets the wavelength sampling
    s2 = _solver(4.0, 10.0, 1e-3)
    ok, msg = s2.check_stability(500.0, 400.0, 10.0)
    assert ok and "wavelength: 4.0 " in msg
    assert f"{500.0 * 1e-3 * math.sqrt(1 / 16 + 1 / 100):.3f}" in msg
