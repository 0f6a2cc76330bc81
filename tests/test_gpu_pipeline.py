"""The whole virtual-source chain on the GPU (preprocessing, correlation
stacked over blocks, Green's functions, F-K filter, resampling, multi-shot
migration) against the same chain built from the oracles (scipy / numpy
restatements and the migration oracle): images within relative L2 1e-4."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _blocks(n_blocks, nt, nch, fs, seed=0):
    """Noise with a slow wave train moving along the cable, at fs."""
    rng = np.random.default_rng(seed)
    out = []
    for b in range(n_blocks):
        x = rng.standard_normal((nt, nch))
        t = np.arange(nt)[:, None] / fs
        pos = np.arange(nch)[None, :] * 8.0
        x += 3.0 * np.sin(2 * np.pi * 2.5 * (t - pos / 900.0) + b)
        out.append(x.astype(np.float32))
    return out


def test_chain_matches_oracle_chain():
    from das_amd.pipeline import migrate_virtual_sources, stacked_correlations, virtual_gathers
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D, ricker_wavelet
    from oracle import interferometry_oracle as IO
    from oracle import oracle as ora
    from oracle import preprocessing_oracle as PO

    fs, target, fmin, fmax, max_lag = 200.0, 50.0, 0.8, 4.5, 1.0
    nch, dx = 48, 8.0
    blocks = _blocks(2, 6000, nch, fs)
    srcs = [10, 20]
    recs = np.arange(24, 46, 2)
    ccfs, lags, dt_proc = stacked_correlations(blocks, fs, target, fmin, fmax, srcs, recs, max_lag)
    # oracle chain
    stack = None
    for raw in blocks:
        dec, dtp, fsp = PO.decimate_time(raw, fs, target)
        w = PO.spectral_whitening(PO.temporal_normalization(PO.bandpass_filter(dec, fsp, fmin,
                                                                              fmax)), fsp, fmin,
                                  fmax)
        c = np.stack([IO.cross_correlation(w, s, max_lag, dtp, recs)[0] for s in srcs])
        stack = c if stack is None else stack + c
    stack = stack / np.float32(len(blocks))
    e = np.linalg.norm(ccfs - stack) / np.linalg.norm(stack)
    print(f"stacked correlations rel L2 {e:.2e}")
    assert e < 1e-4
    coords = np.arange(nch) * dx
    gathers = virtual_gathers(ccfs, lags, srcs, recs, coords, 2 * dx, 300.0, 3500.0)
    want = []
    for k in range(len(srcs)):
        g, _ = IO.branch(IO.time_derivative(stack[k], float(lags[1] - lags[0])), lags)
        want.append(PO.fk_filter(g, float(lags[1] - lags[0]), 2 * dx, 300.0, 3500.0))
    for a, b in zip(gathers, want):
        assert np.linalg.norm(a - b) / np.linalg.norm(b) < 1e-4
    # migration on a small model of the cable's section
    nz, nx, dx_sim, dt_sim = 40, 60, 5.0, 1e-3
    vel = np.full((nz, nx), 1500.0, np.float32)
    vel[20:] = 2200.0
    fm = FloodModels(vel, vel, vel, 0, 0)
    solver = WaveSolver2D(nx, nz, dx_sim, dx_sim, dt_sim, n_sponge=10)
    six = [s // 2 + 10 for s in srcs]
    rix = np.array([r // 2 + 10 for r in recs])
    res = migrate_virtual_sources(TSWRTMEngine(solver), fm, gathers, dt_proc, max_lag, 10.0, six,
                                  rix, 2)
    sim_nt = int(round(max_lag / dt_sim))
    rg = [IO.resample(g, dt_proc, dt_sim, sim_nt) for g in want]
    src, _ = ricker_wavelet(10.0, dt_sim, max_lag)
    g = ora.Grid(nx, nz, dx_sim, dx_sim, dt_sim, 10)
    raw, lap, norm = ora.migrate_multiple(g, vel, vel, [src] * 2, six, [2, 2], rg, rix, 2, 2)
    for got, ref, name in ((res.raw_image, raw, "raw"), (res.laplacian_image, lap, "laplacian")):
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"{name} image rel L2 vs oracle chain {e:.2e}")
        assert e < 1e-4
