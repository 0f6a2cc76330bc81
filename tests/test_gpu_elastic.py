"""Elastic path on the GPU: the reference's elastic step (golden vectors),
the elastic driver loops (golden vectors made with the reference kernel) and
the migration of an elastic shot (200 x 400, nt 1200) against the oracle (FP32 bit-exact, FP64 error
stated)."""
import numpy as np
import pytest

from das_amd import _native as N

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))


@pytest.fixture(scope="module", autouse=True)
def engine():
    N.check(N.lib().tsw_init(0))


def test_elastic_loops_golden(golden):
    from das_amd.wave_solver import WaveSolver2D

    g = golden("elastic_loops")
    for i in range(int(g["ncase"][0])):
        nx, nz, nsp, fs, nt, k, six, siz, riz = (int(v) for v in g[f"c{i}_ints"])
        dx, dt, alpha = (float(v) for v in g[f"c{i}_flt"])
        s = WaveSolver2D(nx, nz, dx, dx, dt, n_sponge=nsp, sponge_alpha=alpha,
                         free_surface_top=bool(fs), solver_type="elastic")
        model = tuple(g[f"c{i}_par"])
        rec, snaps, ill = s.forward_propagate_elastic(model, g[f"c{i}_src"], six, siz,
                                                      g[f"c{i}_rix"], riz, True, k, True)
        np.testing.assert_array_equal(rec, g[f"c{i}_rec"], err_msg=f"rec {i}")
        np.testing.assert_array_equal(snaps, g[f"c{i}_snaps"], err_msg=f"snaps {i}")
        np.testing.assert_array_equal(ill, g[f"c{i}_illum"], err_msg=f"illum {i}")
        img = np.zeros((nz, nx), np.float32)
        s.backward_propagate_and_correlate_elastic(model, g[f"c{i}_data"], g[f"c{i}_rix"], riz,
                                                   snaps, img, k)
        np.testing.assert_array_equal(img, g[f"c{i}_image"], err_msg=f"image {i}")


def test_elastic_migrate_vs_oracle():
    """An elastic shot of the 400 x 1600 workload's kind, reduced to run the CPU oracle quickly
    (200 x 400, nt 1200): FP32 bit-exact to the oracle, FP64 error stated."""
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import PlainElasticShot
    from oracle import oracle as ora

    cfg = PlainElasticShot(nz=200, nx=400, nt=1200, n_rec=200)
    s = WaveSolver2D(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge,
                     solver_type="elastic")
    rec, _ = s.forward_propagate_elastic(cfg.model_true, cfg.src, cfg.six, cfg.siz, cfg.rix,
                                         cfg.riz)
    g = ora.Grid(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, cfg.n_sponge)
    rec_o, _, _ = ora.elastic_forward(g, cfg.model_true, cfg.src, cfg.six, cfg.siz, cfg.rix,
                                      cfg.riz)
    np.testing.assert_array_equal(rec, rec_o)
    img, ill = TSWRTMEngine(s).migrate_shot_elastic(cfg.model_mig, cfg.model_mig, cfg.src,
                                                    cfg.six, cfg.siz, rec, cfg.rix, cfg.riz,
                                                    cfg.k)
    img_o, ill_o = ora.elastic_migrate_shot(g, cfg.model_mig, cfg.model_mig, cfg.src, cfg.six,
                                            cfg.siz, rec_o, cfg.rix, cfg.riz, cfg.k)
    img_d, ill_d = ora.elastic_migrate_shot(g, cfg.model_mig, cfg.model_mig, cfg.src, cfg.six,
                                            cfg.siz, rec_o, cfg.rix, cfg.riz, cfg.k,
                                            dtype=np.float64)
    print(f"\nelastic rel L2 vs FP32 oracle: image {rel_l2(img, img_o):.2e} "
          f"illum {rel_l2(ill, ill_o):.2e}; vs FP64: image {rel_l2(img, img_d):.2e} "
          f"(CPU FP32 {rel_l2(img_o, img_d):.2e})")
    np.testing.assert_array_equal(img, img_o)
    np.testing.assert_array_equal(ill, ill_o)
    assert np.abs(img).max() > 0


@pytest.mark.parametrize("fs", [True, False])
def test_batched_elastic_equals_single_shots(fs):
    """The batched elastic lane advances its shots with one fused launch per
    time step (LDS tiles with a 56 x 24 core); every shot equals its
    single-shot migration, which runs the two-pass per-step kernels, bit for
    bit.  Grids of one tile, of several and with partial last tiles in x and
    z; source and receiver rows 0..3; receivers on edge columns and
    duplicated; odd and even nt; batched, plain and kernel-timed runs; one
    slot per shot and shared slots, with and without hipGraphs."""
    from das_amd.batch import ShotBatch
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.wave_solver import WaveSolver2D
    from das_amd.workloads import ElasticShot, elastic_from_vp, layered_faulted_vp, smooth

    dx, dt, nsp, k = 5.0, 4e-4, 8, 2
    rng = np.random.default_rng(23 + int(fs))
    specs = [  # nz, nx, nt, six, siz, riz
        (23, 55, 31, 20, 0, 0), (25, 57, 40, 56, 1, 1), (49, 113, 37, 3, 3, 2),
        (6, 130, 26, 129, 2, 3), (70, 168, 44, 84, 2, 2)]
    shots, singles = [], []
    for nz, nx, nt, six, siz, riz in specs:
        vp = layered_faulted_vp(nz, nx, dx, seed=int(rng.integers(50)))
        mt = elastic_from_vp(vp)
        mm = elastic_from_vp(smooth(vp, 40.0, dx))
        src = (rng.standard_normal(nt) * 1e3).astype(np.float32)
        rix = np.concatenate([[0, 1, nx - 1, nx - 2, 3, 3],
                              np.arange(2, nx - 2, 4)]).astype(np.int32)
        solver = WaveSolver2D(nx, nz, dx, dx, dt, n_sponge=nsp, free_surface_top=fs,
                              solver_type="elastic")
        data, _ = solver.forward_propagate_elastic(mt, src, six, siz, rix, riz)
        shots.append(ElasticShot(nx=nx, nz=nz, nt=nt, model_fwd=mm, model_bwd=mt, src=src,
                                 six=six, siz=siz, data=data, rix=rix, riz=riz))
        singles.append(TSWRTMEngine(solver).migrate_shot_elastic(mm, mt, src, six, siz, data,
                                                                 rix, riz, k))
    # without the free surface rows 0 and 1 are never updated, so gathers on
    # them are zero and so are those two shots' images
    assert sum(bool(np.abs(img).max() > 0) for img, _ in singles) >= (5 if fs else 3)
    for slots, gs in ((0, 100), (2, 0), (1, 7)):
        b = ShotBatch(shots, dx, dx, dt, n_sponge=nsp, free_surface_top=fs, k=k, slots=slots,
                      graph_steps=gs)
        for mode in ("batched", "plain", "timed"):
            if mode == "timed":
                ms, nl = b.time_kernels("elastic")
                assert nl > 0 and ms > 0
            else:
                b.run(mode)
            for i, (img, ill) in enumerate(singles):
                bi, bl = b.shot_result(i)
                np.testing.assert_array_equal(bi, img, err_msg=f"slots {slots} {mode} shot {i}")
                np.testing.assert_array_equal(bl, ill, err_msg=f"slots {slots} {mode} shot {i}")
        b.close()
