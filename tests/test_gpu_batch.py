"""Batched multi-shot RTM: every shot of a mixed batch equals the same shot
migrated alone (bit for bit), for the per-step kernel and every temporally
blocked configuration, in batched and plain modes and for several slot
counts; the window stack equals the shot-order FP32 sum."""
import numpy as np
import pytest

from das_amd import _native as N

pytestmark = pytest.mark.gpu

DX, DT, NSP, K = 5.0, 5e-4, 8, 2


def _shots(seed=11):
    """Uneven shots: several tiles in x and z, nt not a multiple of T,
    receivers on edge columns and duplicated, source/receiver rows 0..3."""
    from das_amd.workloads import Shot, correlogram_gather, faulted_two_layer, smooth
    from das_amd.wave_solver import ricker_wavelet

    rng = np.random.default_rng(seed)
    specs = [  # nz, nx, nt, six, siz, riz
        (70, 300, 97, 150, 2, 2), (40, 48, 61, 10, 0, 1), (130, 190, 140, 5, 3, 0),
        (55, 131, 77, 130, 1, 3), (64, 256, 120, 128, 2, 2)]
    shots, x0 = [], 0
    for nz, nx, nt, six, siz, riz in specs:
        v = smooth(faulted_two_layer(nz, nx, DX, seed=int(rng.integers(100))), 30.0, DX)
        vb = smooth(faulted_two_layer(nz, nx, DX, seed=int(rng.integers(100))), 50.0, DX)
        rix = np.concatenate([[0, 1, nx - 1, nx - 2, 3, 3], np.arange(2, nx - 2, 3)]).astype(np.int32)
        src, _ = ricker_wavelet(30.0, DT, nt * DT)
        data = correlogram_gather(rng, nt, DT, (rix - six) * DX, f0=30.0, f_lo=5, f_hi=60)
        shots.append(Shot(nx=nx, nz=nz, nt=nt, vel_fwd=v, vel_bwd=vb, src=src, six=six,
                          siz=siz, data=data, rix=rix, riz=riz, x0=x0))
        x0 += nx // 2
    gnx = max(s.x0 + s.nx for s in shots)
    gnz = max(s.nz for s in shots)
    return shots, gnx, gnz


def _singles(shots, fs):
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D

    out = []
    for s in shots:
        solver = WaveSolver2D(s.nx, s.nz, DX, DX, DT, n_sponge=NSP, free_surface_top=fs)
        fm = FloodModels(s.vel_fwd, s.vel_fwd, s.vel_bwd, 0, 0)
        out.append(TSWRTMEngine(solver).migrate_shot(fm, s.src, s.six, s.siz, s.data, s.rix,
                                                     s.riz, K))
    return out


@pytest.fixture(scope="module")
def mixed():
    N.check(N.lib().tsw_init(0))
    shots, gnx, gnz = _shots()
    return shots, gnx, gnz, {True: _singles(shots, True), False: _singles(shots, False)}


def _check(b, shots, singles, label):
    for i, (img, ill) in enumerate(singles):
        bi, bl = b.shot_result(i)
        np.testing.assert_array_equal(bi, img, err_msg=f"{label} shot {i} image")
        np.testing.assert_array_equal(bl, ill, err_msg=f"{label} shot {i} illumination")


def test_single_path_matches_oracle(mixed):
    from oracle import oracle as ora

    shots, _, _, singles = mixed
    for fs in (True, False):
        s = shots[1]
        g = ora.Grid(s.nx, s.nz, DX, DX, DT, NSP, free_surface=fs)
        img, ill = ora.migrate_shot(g, s.vel_fwd, s.vel_bwd, s.src, s.six, s.siz, s.data, s.rix,
                                    s.riz, K)
        np.testing.assert_array_equal(singles[fs][1][0], img)
        np.testing.assert_array_equal(singles[fs][1][1], ill)


@pytest.mark.parametrize("slots,graph_steps", [(0, 100), (2, 7), (1, 0)])
def test_per_step_batch_equals_single_shots(mixed, slots, graph_steps):
    from das_amd.batch import ShotBatch

    shots, gnx, gnz, singles = mixed
    b = ShotBatch(shots, DX, DX, DT, n_sponge=NSP, k=K, slots=slots, graph_steps=graph_steps)
    for mode in ("batched", "plain", "batched"):  # runs are repeatable
        b.run(mode)
        _check(b, shots, singles[True], f"slots={slots} {mode}")
    assert b.info()["cell_updates"] == sum(2 * s.nt * s.nx * s.nz for s in shots)
    gi, gl = b.stack(gnx, gnz)
    ref_i = np.zeros((gnz, gnx), np.float32)
    ref_l = np.zeros_like(ref_i)
    for s, (img, ill) in zip(shots, singles[True]):
        ref_i[:s.nz, s.x0:s.x0 + s.nx] += img
        ref_l[:s.nz, s.x0:s.x0 + s.nx] += ill
    np.testing.assert_array_equal(gi, ref_i)
    np.testing.assert_array_equal(gl, ref_l)
    b.close()


def test_blocked_kernels_equal_single_shots(mixed):
    from das_amd.batch import ShotBatch, kernel_configs

    shots, _, _, singles = mixed
    cfgs = kernel_configs()
    assert len(cfgs) >= 4
    for kid, (T, W, H, threads) in cfgs.items():
        for fs in (True, False):
            for slots in (0, 2):
                print(f"kernel {kid} (T={T}, core {W}x{H}, {threads} thr) fs={fs} slots={slots}",
                      flush=True)
                b = ShotBatch(shots, DX, DX, DT, n_sponge=NSP, k=K, slots=slots, kernel=kid,
                              free_surface_top=fs, graph_steps=100 if slots else 0)
                b.run("batched")
                _check(b, shots, singles[fs], f"kernel {kid} fs={fs} slots={slots}")
                b.run("batched")  # repeat: state regenerated, same result
                _check(b, shots, singles[fs], f"kernel {kid} fs={fs} slots={slots} rerun")
                ms, nl = b.time_kernels()
                assert nl > 0 and ms > 0
                _check(b, shots, singles[fs], f"kernel {kid} timed run")
                b.close()


@pytest.mark.parametrize("groups,graph_steps", [(1, 100), (3, 5), (8, 0), (8, 20000)])
def test_launch_groups_equal_single_shots(mixed, monkeypatch, groups, graph_steps):
    """The blocked kernel's launches split into launch groups on their own
    streams (TSW_LAUNCH_GROUPS, at most one group per slot): every group
    count, with no graphs, one launch per graph or the whole run in one graph,
    gives the single-shot results, and a timed run counts one launch per
    block of T steps whatever the grouping."""
    from das_amd.batch import ShotBatch, kernel_configs

    shots, _, _, singles = mixed
    kid = 25
    T = kernel_configs()[kid][0]
    monkeypatch.setenv("TSW_LAUNCH_GROUPS", str(groups))
    b = ShotBatch(shots, DX, DX, DT, n_sponge=NSP, k=K, slots=0, kernel=kid,
                  graph_steps=graph_steps)
    for _ in range(2):
        b.run("batched")
        _check(b, shots, singles[True], f"groups={groups} graph_steps={graph_steps}")
    ms, nl = b.time_kernels()
    assert ms > 0 and nl == max(2 * -(-s.nt // T) for s in shots)
    _check(b, shots, singles[True], f"groups={groups} timed run")
    b.close()


def test_mixed_acoustic_elastic_batch(mixed):
    """Elastic shots form their own lane; acoustic and elastic shots of one
    batch each equal their single-shot migration bit for bit."""
    from das_amd.batch import ShotBatch
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import ElasticShot, elastic_from_vp, layered_faulted_vp, smooth

    shots, gnx, gnz, singles = mixed
    rng = np.random.default_rng(5)
    eshots, esingle = [], []
    for nz, nx, nt, six, riz in ((40, 70, 50, 30, 2), (56, 100, 41, 60, 1)):
        vp = layered_faulted_vp(nz, nx, DX, seed=int(rng.integers(50)))
        mt = elastic_from_vp(vp)
        mm = elastic_from_vp(smooth(vp, 40.0, DX))
        src = (rng.standard_normal(nt) * 1e3).astype(np.float32)
        rix = np.arange(1, nx - 1, 3).astype(np.int32)
        solver = WaveSolver2D(nx, nz, DX, DX, 4e-4, n_sponge=NSP, solver_type="elastic")
        data, _ = solver.forward_propagate_elastic(mt, src, six, 2, rix, riz)
        eshots.append(ElasticShot(nx=nx, nz=nz, nt=nt, model_fwd=mm, model_bwd=mt, src=src,
                                  six=six, siz=2, data=data, rix=rix, riz=riz, x0=3))
        esingle.append(TSWRTMEngine(solver).migrate_shot_elastic(mm, mt, src, six, 2, data, rix,
                                                                  riz, K))
    # the elastic shots use dt 4e-4; pair them with acoustic shots of the same dt
    from das_amd.workloads import Shot
    ashot = shots[1]
    a = Shot(**{**ashot.__dict__})
    solver = WaveSolver2D(a.nx, a.nz, DX, DX, 4e-4, n_sponge=NSP)
    from das_amd.velocity_flood import FloodModels
    asingle = TSWRTMEngine(solver).migrate_shot(FloodModels(a.vel_fwd, a.vel_fwd, a.vel_bwd, 0, 0),
                                                a.src, a.six, a.siz, a.data, a.rix, a.riz, K)
    order = [eshots[0], a, eshots[1]]
    expect = [esingle[0], asingle, esingle[1]]
    for kernel, slots, gs in ((0, 0, 100), (11, 1, 0), (7, 2, 30)):
        b = ShotBatch(order, DX, DX, 4e-4, n_sponge=NSP, k=K, slots=slots, kernel=kernel,
                      graph_steps=gs)
        for mode in ("batched", "plain"):
            b.run(mode)
            for i, (img, ill) in enumerate(expect):
                bi, bl = b.shot_result(i)
                np.testing.assert_array_equal(bi, img, err_msg=f"k{kernel} {mode} shot {i}")
                np.testing.assert_array_equal(bl, ill, err_msg=f"k{kernel} {mode} shot {i}")
        ms, nl = b.time_kernels("elastic")
        assert nl > 0 and ms > 0
        info = b.info()
        assert info["elastic_cell_updates"] == sum(2 * s.nt * s.nx * s.nz for s in eshots)
        b.close()


def test_full_size_windows_batched_equal_single_shots():
    """Four windows of the 64-shot workload at full size (nz 300, nx
    600-1600, nt 3000-6000): the tuned blocked kernel with every shot resident
    equals migrating each alone, bit for bit."""
    from das_amd.batch import ShotBatch
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import BatchedWindows

    cfg = BatchedWindows(n_shots=4).build(with_data=True)
    b = ShotBatch(cfg.shots, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k, slots=4,
                  kernel=15, graph_steps=100)
    b.run("batched")
    for i, s in enumerate(cfg.shots):
        solver = WaveSolver2D(s.nx, s.nz, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge)
        img, ill = TSWRTMEngine(solver).migrate_shot(
            FloodModels(s.vel_fwd, s.vel_fwd, s.vel_bwd, 0, 0), s.src, s.six, s.siz, s.data,
            s.rix, s.riz, cfg.k)
        bi, bl = b.shot_result(i)
        print(f"shot {i}: {s.nz}x{s.nx}, nt={s.nt}", flush=True)
        np.testing.assert_array_equal(bi, img)
        np.testing.assert_array_equal(bl, ill)
    b.close()
