"""Preprocessing on the GPU against the oracle restatement (the reference's
module is absent: parity to it is unpinned).  Zero-phase filtering,
decimation and 1-bit are bit-equal to scipy.signal.sosfiltfilt / numpy;
whitening and the F-K fan (hipFFT transforms) agree with numpy's FFTs to
FP32 transform rounding and with the float64 restatement."""
import numpy as np
import pytest

from oracle import preprocessing_oracle as PO

pytestmark = pytest.mark.gpu


def _noise(nt, nch, seed=0, scale=1.0):
    return (np.random.default_rng(seed).standard_normal((nt, nch)) * scale).astype(np.float32)


@pytest.mark.parametrize("nt,nch", [(3000, 64), (200, 3), (50, 1)])
def test_bandpass_bit_equal_to_scipy(nt, nch):
    from das_amd.preprocessing import bandpass_filter

    x = _noise(nt, nch, seed=nt, scale=37.0)
    np.testing.assert_array_equal(bandpass_filter(x, 50.0, 0.8, 4.5), PO.bandpass_filter(x, 50.0,
                                                                                         0.8, 4.5))


def test_short_input_rejected_like_scipy():
    from das_amd.preprocessing import bandpass_filter

    with pytest.raises(ValueError, match="padlen"):
        bandpass_filter(_noise(20, 2), 50.0, 0.8, 4.5)


@pytest.mark.parametrize("fs,target", [(1000.0, 50.0), (250.0, 50.0), (50.0, 50.0), (100.0, 30.0)])
def test_decimation_bit_equal(fs, target):
    from das_amd.preprocessing import decimate_time

    x = _noise(4003, 17, seed=3)
    y, dt, f2 = decimate_time(x, fs, target)
    yo, dto, f2o = PO.decimate_time(x, fs, target)
    np.testing.assert_array_equal(y, yo)
    assert (dt, f2) == pytest.approx((dto, f2o))


def test_one_bit_bit_equal():
    from das_amd.preprocessing import temporal_normalization

    x = _noise(777, 9, seed=4)
    x[0, :4] = [0.0, -0.0, np.inf, -np.inf]
    np.testing.assert_array_equal(temporal_normalization(x), PO.temporal_normalization(x))
    with pytest.raises(ValueError):
        temporal_normalization(x, method="rms")


@pytest.mark.parametrize("nt", [4096, 3001])
def test_whitening_matches_numpy(nt):
    from das_amd.preprocessing import spectral_whitening

    x = np.sign(_noise(nt, 11, seed=nt))
    got = spectral_whitening(x, 50.0, 0.8, 4.5)
    exact = PO.spectral_whitening(x, 50.0, 0.8, 4.5, work_dtype=np.float64)
    ref32 = PO.spectral_whitening(x, 50.0, 0.8, 4.5)
    e64 = np.linalg.norm(got - exact) / np.linalg.norm(exact)
    e32 = np.linalg.norm(got - ref32) / np.linalg.norm(exact)
    print(f"whitening rel L2: vs float64 {e64:.2e}, vs numpy float32 {e32:.2e}")
    assert e64 < 1e-5 and e32 < 1e-5


@pytest.mark.parametrize("nt,nx", [(512, 128), (351, 95)])
def test_fk_filter_matches_numpy(nt, nx):
    from das_amd.preprocessing import fk_filter

    g = _noise(nt, nx, seed=nx)
    got = fk_filter(g, 0.02, 8.0, 800.0, 3500.0)
    exact = PO.fk_filter(g, 0.02, 8.0, 800.0, 3500.0, work_dtype=np.complex128)
    e = np.linalg.norm(got - exact) / np.linalg.norm(exact)
    print(f"F-K rel L2 vs float64: {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6, 7, 9, 12, 16, 18, 32])
def test_zero_phase_filter_any_section_count(order):
    """The filter kernel for every cascade length (1..8 sections unrolled, the
    rest through the general kernel), low-pass and band-pass designs, equals
    scipy.signal.sosfiltfilt bit for bit, as does its decimating form."""
    from scipy import signal

    from das_amd.preprocessing import _sosfiltfilt

    x = _noise(1500, 5, seed=order, scale=3.0)
    for sos in (signal.butter(order, 0.2, output="sos"),
                signal.butter(max(1, order // 2), [0.05, 0.3], btype="band", output="sos")):
        want = signal.sosfiltfilt(sos, x, axis=0).astype(np.float32)
        np.testing.assert_array_equal(_sosfiltfilt(x, sos), want)
        np.testing.assert_array_equal(_sosfiltfilt(x, sos, step=3), want[::3])


def test_transform_plans_per_shape():
    """Whitening and the F-K fan keep a plan per shape: shapes alternating
    between calls give the same results as each shape alone."""
    from das_amd.preprocessing import fk_filter, spectral_whitening

    a, b = _noise(600, 7, seed=1), _noise(256, 12, seed=2)
    wa, wb = spectral_whitening(a, 50.0, 1.0, 10.0), spectral_whitening(b, 50.0, 1.0, 10.0)
    fa, fb = fk_filter(a, 0.02, 8.0, 100.0, 900.0), fk_filter(b, 0.02, 8.0, 100.0, 900.0)
    for _ in range(2):
        np.testing.assert_array_equal(spectral_whitening(b, 50.0, 1.0, 10.0), wb)
        np.testing.assert_array_equal(spectral_whitening(a, 50.0, 1.0, 10.0), wa)
        np.testing.assert_array_equal(fk_filter(b, 0.02, 8.0, 100.0, 900.0), fb)
        np.testing.assert_array_equal(fk_filter(a, 0.02, 8.0, 100.0, 900.0), fa)


@pytest.mark.parametrize("fs_orig,target,nt,nch", [(100.0, 50.0, 3001, 37), (200.0, 50.0, 6000, 16),
                                                   (50.0, 50.0, 2000, 9)])
def test_device_chain_equals_step_by_step(fs_orig, target, nt, nch):
    """One call that keeps the block on the device (decimation, band-pass
    with the 1-bit sign fused into its output, whitening) equals the four
    host-pointer functions in turn, bit for bit; with and without decimation,
    odd block lengths, a reused output tensor."""
    from das_amd.preprocessing import (bandpass_filter, decimate_time, preprocess_block_device,
                                       spectral_whitening, temporal_normalization)

    out = None
    for seed in (1, 2):
        x = _noise(nt, nch, seed=seed, scale=5.0)
        dec, dt, fs = decimate_time(x, fs_orig, target)
        want = spectral_whitening(temporal_normalization(bandpass_filter(dec, fs, 0.8, 4.5)), fs,
                                  0.8, 4.5)
        got, dt2, fs2 = preprocess_block_device(x, fs_orig, target, 0.8, 4.5, out=out)
        assert (dt2, fs2) == (dt, fs)
        assert out is None or got.data_ptr() == out.data_ptr()
        out = got
        np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_device_chain_rejects_short_blocks_like_scipy():
    from das_amd.preprocessing import preprocess_block_device

    with pytest.raises(ValueError, match="padlen"):
        preprocess_block_device(_noise(40, 2), 100.0, 50.0, 0.8, 4.5)


def test_stacked_correlations_equal_host_chain():
    """The virtual-source stack built on the device (each block copied once,
    the stack resident until the end) equals the host-pointer chain with the
    FP32 stack accumulated on the host, bit for bit."""
    from das_amd.interferometry import cross_correlations
    from das_amd.pipeline import stacked_correlations
    from das_amd.preprocessing import (bandpass_filter, decimate_time, spectral_whitening,
                                       temporal_normalization)

    nch = 40
    blocks = [_noise(4000, nch, seed=s) for s in (4, 5, 6)]
    srcs, recs = [3, -2], np.arange(0, nch, 3)
    ccfs, lags, dtp = stacked_correlations(blocks, 100.0, 50.0, 0.8, 4.5, srcs, recs, 2.0)
    stack = None
    for raw in blocks:
        dec, dt, fs = decimate_time(raw, 100.0, 50.0)
        w = spectral_whitening(temporal_normalization(bandpass_filter(dec, fs, 0.8, 4.5)), fs,
                               0.8, 4.5)
        if stack is None:
            stack, want_lags = cross_correlations(w, srcs, 2.0, dt, recs)
        else:
            cross_correlations(w, srcs, 2.0, dt, recs, out=stack)
    assert dtp == dt
    np.testing.assert_array_equal(lags, want_lags)
    np.testing.assert_array_equal(ccfs, stack / np.float32(len(blocks)))
    assert np.abs(ccfs).max() > 0


@pytest.mark.parametrize("fs_orig,target,nt,nch", [(100.0, 50.0, 3001, 37), (50.0, 50.0, 2000, 70)])
def test_device_blocks_equal_single_blocks(fs_orig, target, nt, nch):
    """Several blocks in one call (their channels filtered side by side, a
    wave straddling two blocks when nch is not a multiple of 64) equal each
    block preprocessed alone, bit for bit; reused outputs."""
    from das_amd.preprocessing import preprocess_block_device, preprocess_blocks_device

    xs = [_noise(nt, nch, seed=s, scale=3.0) for s in (11, 12, 13)]
    want = [preprocess_block_device(x, fs_orig, target, 0.8, 4.5)[0].cpu().numpy() for x in xs]
    outs = None
    for _ in range(2):
        outs, dt, fs = preprocess_blocks_device(xs, fs_orig, target, 0.8, 4.5, outs=outs)
        assert len(outs) == 3
        for o, w in zip(outs, want):
            np.testing.assert_array_equal(o.cpu().numpy(), w)
    with pytest.raises(ValueError, match="one shape"):
        preprocess_blocks_device([xs[0], xs[1][:-1]], fs_orig, target, 0.8, 4.5)


@pytest.mark.parametrize("group", [1, 2, 4])
def test_stacked_correlations_any_group(group):
    """The stack is the same whatever number of blocks are preprocessed per
    call, including a last, smaller group and a change of block shape."""
    from das_amd.pipeline import stacked_correlations

    nch = 40
    blocks = [_noise(4000, nch, seed=s) for s in (4, 5, 6)] + [_noise(3000, nch, seed=7)]
    ref, lags, dtp = stacked_correlations(blocks, 100.0, 50.0, 0.8, 4.5, [3, -2],
                                          np.arange(0, nch, 3), 2.0, group=1)
    got, lags2, dtp2 = stacked_correlations(blocks, 100.0, 50.0, 0.8, 4.5, [3, -2],
                                            np.arange(0, nch, 3), 2.0, group=group)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(lags2, lags)
    assert dtp2 == dtp
