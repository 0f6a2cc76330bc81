"""Divergence sentinel: a shot whose time step violates the CFL limit ends with
NaN/Inf in its image; the engine reports it (status -6, WavefieldDiverged, a
FloatingPointError) instead of returning the image as a success.  Every other
shot of the same batch, stack or serving queue keeps its exact result.

The reference has no such check (its only guard is the advisory
check_stability, tsw_rtm/wave_solver.py:119-126): the oracle, which follows
it, returns the non-finite image, which pins that these inputs do diverge."""
import numpy as np
import pytest

from das_amd import _native as N

pytestmark = pytest.mark.gpu

DX, DT, NSP, K = 5.0, 5e-4, 8, 2
FAST_MUL = 15.0  # ~1500 m/s models x 15: c*dt/dx ~ 2.2, far above the scheme's limit


def _shot(nz, nx, nt, seed, vmul=1.0, x0=0):
    from das_amd.wave_solver import ricker_wavelet
    from das_amd.workloads import Shot, correlogram_gather, faulted_two_layer, smooth

    rng = np.random.default_rng(seed)
    v = smooth(faulted_two_layer(nz, nx, DX, seed=seed), 30.0, DX) * np.float32(vmul)
    rix = np.arange(2, nx - 2, 3).astype(np.int32)
    six = nx // 2
    src, _ = ricker_wavelet(30.0, DT, nt * DT)
    data = correlogram_gather(rng, nt, DT, (rix - six) * DX, f0=30.0, f_lo=5, f_hi=60)
    return Shot(nx=nx, nz=nz, nt=nt, vel_fwd=v, vel_bwd=v, src=src, six=six, siz=2,
                data=data, rix=rix, riz=2, x0=x0)


def _engine(s, custom_sponge=False):
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.wave_solver import WaveSolver2D

    solver = WaveSolver2D(s.nx, s.nz, DX, DX, DT, n_sponge=NSP)
    if custom_sponge:
        solver.sponge = solver.sponge.copy() * np.float32(0.999)
    return TSWRTMEngine(solver)


def _migrate(s, custom_sponge=False):
    from das_amd.velocity_flood import FloodModels

    return _engine(s, custom_sponge).migrate_shot(
        FloodModels(s.vel_fwd, s.vel_fwd, s.vel_bwd, 0, 0), s.src, s.six, s.siz, s.data,
        s.rix, s.riz, K)


def test_oracle_diverges_on_these_inputs():
    from oracle import oracle as ora

    s = _shot(40, 64, 60, 1, vmul=FAST_MUL)
    g = ora.Grid(s.nx, s.nz, DX, DX, DT, NSP, free_surface=True)
    img, ill = ora.migrate_shot(g, s.vel_fwd, s.vel_bwd, s.src, s.six, s.siz, s.data, s.rix,
                                s.riz, K)
    assert not np.isfinite(img).all() and not np.isfinite(ill).all()


@pytest.mark.parametrize("custom_sponge", [False, True])
def test_migrate_shot_raises_on_divergence(custom_sponge):
    N.check(N.lib().tsw_init(0))
    bad = _shot(40, 64, 60, 1, vmul=FAST_MUL)
    with pytest.raises(N.WavefieldDiverged, match="diverged") as ei:
        _migrate(bad, custom_sponge)
    assert isinstance(ei.value, FloatingPointError) and ei.value.status == -6
    img, ill = ei.value.result  # the finished arrays travel with the error
    assert not np.isfinite(img).all() and not np.isfinite(ill).all()
    # a stable shot right after it is unaffected
    good = _shot(40, 64, 60, 1)
    img, ill = _migrate(good, custom_sponge)
    assert np.isfinite(img).all() and np.isfinite(ill).all()


def test_elastic_shot_raises_on_divergence():
    from das_amd.wave_solver import WaveSolver2D, ricker_wavelet
    from das_amd.workloads import elastic_from_vp, layered_faulted_vp

    N.check(N.lib().tsw_init(0))
    nz, nx, nt = 40, 80, 60
    vp = layered_faulted_vp(nz, nx, DX, seed=2) * np.float32(8.0)
    m = elastic_from_vp(vp)
    src, _ = ricker_wavelet(25.0, DT, nt * DT)
    rix = np.arange(2, nx - 2, 2).astype(np.int32)
    data = np.random.default_rng(3).standard_normal((nt, len(rix))).astype(np.float32)
    eng = _engine(_shot(nz, nx, nt, 2))
    eng.solver = WaveSolver2D(nx, nz, DX, DX, DT, n_sponge=NSP, solver_type="elastic")
    with pytest.raises(N.WavefieldDiverged, match="elastic shot"):
        eng.migrate_shot_elastic(m, m, src, nx // 2, 2, data, rix, 2, K)


def test_batch_names_the_diverged_shot_and_keeps_the_others():
    from das_amd.batch import ShotBatch

    N.check(N.lib().tsw_init(0))
    shots = [_shot(48, 128, 90, 4), _shot(40, 96, 70, 5, x0=60),
             _shot(56, 160, 80, 6, vmul=FAST_MUL, x0=100), _shot(48, 110, 75, 7, x0=150)]
    singles = {i: _migrate(s) for i, s in enumerate(shots) if i != 2}
    for slots in (0, 2):
        b = ShotBatch(shots, DX, DX, DT, n_sponge=NSP, k=K, slots=slots,
                      graph_steps=100 if slots else 0)
        with pytest.raises(N.WavefieldDiverged, match="shot 2:"):
            b.run("batched")
        for i, (img, ill) in singles.items():
            bi, bl = b.shot_result(i)
            np.testing.assert_array_equal(bi, img, err_msg=f"slots {slots} shot {i}")
            np.testing.assert_array_equal(bl, ill, err_msg=f"slots {slots} shot {i}")
        assert not np.isfinite(b.shot_result(2)[0]).all()
        b.close()


def test_multi_shot_stack_raises_with_the_outputs():
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D

    N.check(N.lib().tsw_init(0))
    s = _shot(40, 96, 60, 8, vmul=FAST_MUL)
    eng = TSWRTMEngine(WaveSolver2D(s.nx, s.nz, DX, DX, DT, n_sponge=NSP))
    fm = FloodModels(s.vel_fwd, s.vel_fwd, s.vel_bwd, 0, 0)
    # one model for both shots: both diverge, the first is named
    with pytest.raises(N.WavefieldDiverged, match="shot 0: .*diverged") as ei:
        eng.migrate_multiple_shots(fm, [s.src] * 2, [10, s.six], [2, 2], [s.data] * 2, s.rix,
                                   2, K)
    raw, lap, norm = ei.value.result
    assert not np.isfinite(raw).all()


def test_served_request_fails_with_the_reason_and_others_finish():
    from das_amd.serve import RTMServer

    N.check(N.lib().tsw_init(0))
    shots = [_shot(48, 128, 90, 9), _shot(56, 160, 80, 10, vmul=FAST_MUL),
             _shot(40, 96, 70, 11)]
    expect = {0: _migrate(shots[0]), 2: _migrate(shots[2])}
    srv = RTMServer(DX, DX, DT, n_sponge=NSP, k=K, budget_bytes=2e9, max_active=3,
                    kernel=11, tick_steps=40)
    ids = [srv.submit(s) for s in shots]
    polls = [srv.wait(r) for r in ids]
    assert [p["state"] for p in polls] == ["done", "failed", "done"]
    with pytest.raises(N.TswError, match="diverged"):
        srv.result(ids[1])
    for i in (0, 2):
        img, ill = srv.result(ids[i])
        np.testing.assert_array_equal(img, expect[i][0])
        np.testing.assert_array_equal(ill, expect[i][1])
    st = srv.stats()
    assert st["completed"] == 2 and st["failed"] == 1 and st["mem_used"] == 0
    srv.close()
