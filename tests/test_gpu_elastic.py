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
