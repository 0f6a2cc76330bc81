"""Virtual-shot gather construction on the GPU against the oracle: the
correlation equals the exact lag sum (float64 brute force) to one FP32
rounding and the reference's FFT route to its rounding; stacking over files
is the FP32 running sum; the derivative, branches and resampling are bit-equal
to numpy / scipy."""
import numpy as np
import pytest

from oracle import interferometry_oracle as IO

pytestmark = pytest.mark.gpu


def _data(nt, nch, seed=0):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nt, nch)).astype(np.float32)
    return np.sign(x).astype(np.float32) * np.abs(x) ** 0.5  # uneven amplitudes


@pytest.mark.parametrize("nt,nch,nlag", [(1500, 40, 175), (4096, 9, 600), (300, 5, 0),
                                         (200, 7, 255), (1, 3, 0)])
def test_correlation_matches_oracle(nt, nch, nlag):
    from das_amd.interferometry import compute_cross_correlation, cross_correlations

    x = _data(nt, nch, seed=nt)
    dt = 0.02
    rec = np.arange(1, nch, 2)
    srcs = [0, nch // 2, -1]
    many, lags = cross_correlations(x, srcs, nlag * dt, dt, rec)
    assert many.shape == (3, 2 * nlag + 1, len(rec)) and many.dtype == np.float32
    for k, s in enumerate(srcs):
        exact = IO.direct_correlation(x, s, nlag, rec)
        # one rounding of the exact sum: within 1 ulp of float32
        np.testing.assert_allclose(many[k], exact.astype(np.float32), rtol=1.2e-7,
                                   atol=1e-30)
        ref, ref_lags = IO.cross_correlation(x, s, nlag * dt, dt, rec)
        np.testing.assert_array_equal(lags, ref_lags)
        if np.linalg.norm(exact):
            err = np.linalg.norm(ref - many[k]) / np.linalg.norm(exact)
            assert err < 1e-5, err
        single, _ = compute_cross_correlation(x, s, nlag * dt, dt, rec)
        np.testing.assert_array_equal(single, many[k])


def test_stacking_over_files_is_fp32_running_sum():
    from das_amd.interferometry import cross_correlations

    files = [_data(800, 12, seed=s) for s in range(3)]
    rec = np.arange(0, 12, 3)
    out = np.zeros((2, 161, len(rec)), np.float32)
    want = np.zeros_like(out)
    for x in files:
        c, _ = cross_correlations(x, [2, 5], 0.8, 0.01, rec)
        want = want + c
        cross_correlations(x, [2, 5], 0.8, 0.01, rec, out=out)
    np.testing.assert_array_equal(out, want)


@pytest.mark.parametrize("nl,ncol", [(351, 40), (2, 3), (64, 5), (1001, 7)])
def test_time_derivative(nl, ncol):
    from das_amd.interferometry import time_derivative_greens

    c = _data(nl, ncol, seed=nl)
    lags = np.arange(-(nl // 2), nl - nl // 2) * 0.02
    dt = float(lags[1] - lags[0])  # as build_virtual_shot_gather forms it
    np.testing.assert_array_equal(time_derivative_greens(c, dt), IO.time_derivative(c, dt))
    sp = time_derivative_greens(c, dt, scheme="spectral")
    exact = IO.time_derivative(c, dt, scheme="spectral", work_dtype=np.float64)
    ref32 = IO.time_derivative(c, dt, scheme="spectral")
    if not np.abs(exact).max():  # two samples: only the dropped Nyquist term
        assert not np.abs(sp).max()
        return
    np.testing.assert_allclose(sp, exact, rtol=1e-5, atol=1e-6 * np.abs(exact).max())
    assert np.linalg.norm(sp - ref32) / np.linalg.norm(exact) < 1e-5


@pytest.mark.parametrize("branch", ["causal", "anticausal", "symmetric"])
@pytest.mark.parametrize("take_derivative", [True, False])
def test_virtual_shot_gather(branch, take_derivative):
    from das_amd.interferometry import build_virtual_shot_gather, compute_cross_correlation

    x = _data(2000, 30, seed=5)
    rec = np.arange(4, 30, 4)
    dt = 0.02
    ccf, lags = compute_cross_correlation(x, 3, 3.5, dt, rec)
    coords = np.arange(30) * 8.0
    g = build_virtual_shot_gather(ccf, lags, 1, 3, coords, rec, branch=branch,
                                  take_derivative=take_derivative)
    d = IO.time_derivative(ccf, float(lags[1] - lags[0])) if take_derivative else ccf
    want, t = IO.branch(d, lags, branch)
    np.testing.assert_array_equal(g.shot_gather, want)
    np.testing.assert_array_equal(g.time_axis, t)
    assert g.dt == float(lags[1] - lags[0])
    np.testing.assert_array_equal(g.offsets, np.abs(coords[rec] - coords[3]))


@pytest.mark.parametrize("nt_o,dt_o,dt_s,target", [(176, 0.02, 0.0015, 2333), (50, 0.004, 0.004, 50),
                                                   (10, 0.1, 0.03, 100), (2, 1.0, 0.25, 9)])
def test_resample_matches_scipy(nt_o, dt_o, dt_s, target):
    from das_amd.interferometry import resample_gather_to_sim_dt

    g = _data(nt_o, 13, seed=nt_o)
    np.testing.assert_array_equal(resample_gather_to_sim_dt(g, dt_o, dt_s, target),
                                  IO.resample(g, dt_o, dt_s, target))


def test_argument_errors():
    from das_amd import _native as N
    from das_amd.interferometry import compute_cross_correlation, resample_gather_to_sim_dt

    x = _data(100, 4)
    with pytest.raises(IndexError):
        compute_cross_correlation(x, 4, 0.1, 0.01)
    with pytest.raises(ValueError, match="correlation window"):
        compute_cross_correlation(x, 0, 1.28, 0.01)  # 128 lags > 2^8/2 - 1
    with pytest.raises(ValueError):
        resample_gather_to_sim_dt(x[:1], 0.01, 0.01, 5)
    assert N.lib().tsw_xcorr(N.ptr(x), 100, 4, None, 1, None, 1, 3, 0, N.ptr(x)) == -1
