"""Host-side argument handling of the correlation and preprocessing calls
(no GPU): channel lists and lag counts as interferometry.py:26-79 forms them,
the zero-phase filter's pad length and initial conditions as
scipy.signal.sosfiltfilt forms them, and the engine loaded after torch so
that it binds torch's HIP runtime."""
import numpy as np
import pytest
from scipy import signal


def test_channel_lists_and_lags():
    from das_amd.interferometry import _xcorr_args

    src, rec, nlag, shape = _xcorr_args(10, [0, -1, 3], None, 0.25, 0.02)
    np.testing.assert_array_equal(src, [0, 9, 3])
    np.testing.assert_array_equal(rec, np.arange(10))
    assert src.dtype == np.int32 and rec.dtype == np.int32
    assert src.flags.c_contiguous and rec.flags.c_contiguous
    assert nlag == 12 and shape == (3, 25, 10)
    src, rec, _, shape = _xcorr_args(10, 4, np.array([-2, 0, 5]), 0.1, 0.02)
    np.testing.assert_array_equal(rec, [8, 0, 5])
    assert shape[0] == 1 and shape[2] == 3
    for bad in ([10], [-11]):
        with pytest.raises(IndexError, match="source_channel_idx"):
            _xcorr_args(10, bad, None, 0.1, 0.02)
    with pytest.raises(IndexError, match="receiver_indices"):
        _xcorr_args(10, [0], np.array([0, 10]), 0.1, 0.02)


@pytest.mark.parametrize("order,band", [(8, None), (4, (0.8, 4.5)), (3, None)])
def test_zero_phase_filter_arguments_as_scipy(order, band):
    from das_amd.preprocessing import _sos_args

    if band is None:
        sos = signal.butter(order, 25.0, btype="lowpass", fs=100.0, output="sos")
    else:
        sos = signal.butter(order, band, btype="bandpass", fs=50.0, output="sos")
    s, zi, nsec, padlen = _sos_args(sos, 1000)
    assert nsec == sos.shape[0] and s.dtype == np.float64 and s.flags.c_contiguous
    np.testing.assert_array_equal(zi, signal.sosfilt_zi(sos))
    # scipy.signal.sosfiltfilt's default: 3 * (2 * n_sections + 1 - trailing zeros)
    x = np.random.default_rng(0).standard_normal(padlen + 1)
    signal.sosfiltfilt(sos, x)  # accepted at padlen + 1 samples
    with pytest.raises(ValueError):
        signal.sosfiltfilt(sos, x[:padlen])
    with pytest.raises(ValueError, match="padlen"):
        _sos_args(sos, padlen)


def test_engine_binds_torch_hip_runtime():
    torch = pytest.importorskip("torch")
    import ctypes
    import os

    from das_amd import _native as N

    if not os.path.exists(N.LIB_PATH):
        pytest.skip("engine not built")
    N.lib()
    hip = os.path.join(os.path.dirname(torch.__file__), "lib", "libamdhip64.so")
    if not os.path.exists(hip):
        pytest.skip("this torch build does not ship its own HIP runtime")
    # the engine's hipGetDevice is torch's
    mine = ctypes.cast(N.lib().hipGetDevice, ctypes.c_void_p).value
    theirs = ctypes.cast(ctypes.CDLL(hip).hipGetDevice, ctypes.c_void_p).value
    assert mine == theirs


def test_streaming_kernels_refuse_wide_sponges():
    """The z-streaming blocked configurations keep the sponge table in LDS
    (at most 255 cells); a wider strip is refused with a bad-argument status
    before anything touches a device, while the tile configurations accept
    it."""
    import ctypes

    from das_amd import _native as N

    L = N.lib()
    ts = [i for i in range(1, L.tsw_kernel_configs() + 1)
          if L.tsw_kernel_config(i, None, None, None, None) == 0 and _wz1(L, i)]
    assert ts, "no streaming configuration in the table"
    d = (N.ShotDesc * 1)()
    for kid in ts:
        h = N.P()
        rc = L.tsw_batch_create(1, d, 5.0, 5.0, 5e-4, 256, 0.0015, 1, 2, 1, 0, kid,
                                ctypes.byref(h))
        assert rc == -1 and "sponge" in N.last_error(), (kid, N.last_error())


def _wz1(L, i):
    """Configuration i is a streaming one: its workgroup has one row of
    threads across the strip (core + 4T halo columns) per sub-step plus one
    that writes back, T + 1 rows, where no tile configuration has T + 1 rows
    of threads."""
    import ctypes

    v = [ctypes.c_int() for _ in range(4)]
    L.tsw_kernel_config(i, *[ctypes.byref(x) for x in v])
    T, W, H, threads = (x.value for x in v)
    return (W + 4 * T) * (T + 1) == threads
