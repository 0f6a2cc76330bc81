"""Pins the CPU oracle (oracle/) to the reference.

Golden vectors (tests/golden/*.npz) come from wave2d.c compiled from source
and from the reference's numpy/scipy expressions (tests/golden/gen_golden.py).
Invariants cover what the vectors leave open: closed forms of the stencil
symbol and the Ricker wavelet, exact linearity, mirror symmetry, and the
FP64 instantiation against FP32.
"""
import numpy as np
import pytest

from oracle import oracle as ora


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def test_sponge_golden(golden):
    g = golden("sponge")
    for i in range(int(g["ncase"][0])):
        nx, nz, n, fs = (int(v) for v in g[f"c{i}_args"])
        a = float(g[f"c{i}_alpha"][0])
        out = ora.build_sponge(nx, nz, n, np.float32(a), bool(fs))
        np.testing.assert_array_equal(out, g[f"c{i}_out"])


def test_sponge_mute_threshold():
    # sponge < 0.99 selects depth >= 3 at alpha = 0.0015 (rtm_engine.py:147)
    sp = ora.build_sponge(100, 80, 30, 0.0015, True)
    row = sp[30, :]
    assert row[30] == 1.0 and row[29] > 0.99 and row[28] > 0.99 and row[27] < 0.99
    assert np.all(sp[0, 31:69] == 1.0)  # no top strip under a free surface
    sp0 = ora.build_sponge(100, 80, 30, 0.0015, False)
    assert sp0[0, 50] < 0.99


def test_acoustic_step_golden(golden):
    g = golden("acoustic_step")
    for i in range(int(g["ncase"][0])):
        pc, pp, pn, vel = g[f"c{i}_in"]
        sp = g[f"c{i}_sponge"]
        sp = None if sp.size == 0 else sp
        dx, dz, dt = (np.float32(v) for v in g[f"c{i}_par"])
        fs = int(g[f"c{i}_fs"][0])
        out = pn.copy()
        ora.acoustic_step(pc, pp, out, vel, sp, dx, dz, dt, fs)
        np.testing.assert_array_equal(out, g[f"c{i}_out"], err_msg=f"case {i}")


def test_elastic_step_golden(golden):
    g = golden("elastic_step")
    for i in range(int(g["ncase"][0])):
        fields = [f.copy() for f in g[f"c{i}_init"]]
        lam, mu, bx, bz = g[f"c{i}_par_fields"]
        sp = g[f"c{i}_sponge"]
        sp = None if sp.size == 0 else sp
        dx, dz, dt = (np.float32(v) for v in g[f"c{i}_par"])
        fs, nsteps = (int(v) for v in g[f"c{i}_ints"])
        for _ in range(nsteps):
            ora.elastic_step(fields, lam, mu, bx, bz, sp, dx, dz, dt, fs)
        for j, name in enumerate(("vx", "vz", "sxx", "szz", "sxz")):
            np.testing.assert_array_equal(fields[j], g[f"c{i}_out"][j], err_msg=f"case {i} {name}")


def test_correlate_golden(golden):
    g = golden("correlate")
    for i in range(int(g["ncase"][0])):
        a, b, img = g[f"c{i}_in"]
        img = img.copy()
        ora.correlate(a, b, img, np.float32(g[f"c{i}_w"][0]))
        np.testing.assert_array_equal(img, g[f"c{i}_out"])


def _grid_from(ints, flt):
    nx, nz, nsp, fs = (int(v) for v in ints[:4])
    return ora.Grid(nx=nx, nz=nz, dx=float(flt[0]), dz=float(flt[0]), dt=float(flt[1]),
                    n_sponge=nsp, sponge_alpha=float(flt[2]), free_surface=bool(fs))


def test_loops_golden(golden):
    g = golden("loops")
    for i in range(int(g["ncase"][0])):
        ints = g[f"c{i}_ints"]
        grid = _grid_from(ints, g[f"c{i}_flt"])
        nt, k, six, siz, riz = (int(v) for v in ints[4:])
        rix = g[f"c{i}_rix"]
        vel_t, vel_m = g[f"c{i}_vel"]
        rec, snaps = ora.forward(grid, vel_t, g[f"c{i}_src"], six, siz, rix, riz, k, save=True)
        np.testing.assert_array_equal(rec, g[f"c{i}_rec"], err_msg=f"rec case {i}")
        np.testing.assert_array_equal(snaps, g[f"c{i}_snaps"], err_msg=f"snaps case {i}")
        img = g[f"c{i}_img0"].copy()
        ora.backward(grid, vel_m, g[f"c{i}_data"], rix, riz, snaps, img, k)
        np.testing.assert_array_equal(img, g[f"c{i}_img"], err_msg=f"image case {i}")


def test_backward_short_snapshot_list(golden):
    # cursor starts past the list end and never moves: image untouched
    g = golden("loops")
    assert not g["short_img"].any()
    grid = ora.Grid(16, 12, 5.0, 5.0, 5e-4, 2, 0.03, True)
    img = np.zeros((12, 16), np.float32)
    rng = np.random.default_rng(0)
    ora.backward(grid, np.full((12, 16), 2000, np.float32), rng.standard_normal((20, 2)),
                 np.array([3, 7]), 2, rng.standard_normal((3, 12, 16)).astype(np.float32), img, 2)
    assert not img.any()


def _mig_setup(g):
    nx, nz, nsp, nt, k, siz, riz = (int(v) for v in g["ints"])
    dx, dz, dt, alpha = (float(v) for v in g["flt"])
    grid = ora.Grid(nx, nz, dx, dz, dt, nsp, alpha, True)
    return grid, nt, k, siz, riz


def test_migrate_golden(golden):
    g = golden("migrate")
    grid, nt, k, siz, riz = _mig_setup(g)
    _, vel_f, vel_b = g["vel"]
    stacked = np.zeros((grid.nz, grid.nx), np.float32)
    illum = np.zeros_like(stacked)
    for j, six in enumerate(g["sixs"]):
        img, ill = ora.migrate_shot(grid, vel_f, vel_b, g["src"], int(six), siz, g[f"s{j}_data"],
                                    g["rix"], riz, k)
        np.testing.assert_array_equal(img, g[f"s{j}_img"])
        np.testing.assert_array_equal(ill, g[f"s{j}_ill"])
        stacked += img
        illum += ill
    np.testing.assert_array_equal(stacked, g["stacked"])
    for mute, sfx in ((True, ""), (False, "_nm")):
        raw, lap, norm = ora.postprocess(grid, stacked, illum, mute)
        np.testing.assert_array_equal(raw, g["raw" + sfx])
        np.testing.assert_array_equal(norm, g["norm" + sfx])
        np.testing.assert_array_equal(lap, g["lap" + sfx])


def test_laplacian_golden(golden):
    g = golden("migrate")
    for i in range(4):
        dx, dz = g[f"lap{i}_d"]
        out = ora.laplacian(g[f"lap{i}_in"], float(dx), float(dz))
        ref = g[f"lap{i}_out"]
        assert out.shape == ref.shape
        np.testing.assert_array_equal(out, ref)


def test_migrate_reference_kernels_equal_restatement():
    if ora.ref_lib() is None:
        pytest.skip("reference engine not compiled here")
    rng = np.random.default_rng(3)
    grid = ora.Grid(64, 40, 5.0, 5.0, 5e-4, 8)
    vel = rng.uniform(1500, 2500, (40, 64)).astype(np.float32)
    src, _ = ora.ricker(25.0, 5e-4, 0.05)
    data = rng.standard_normal((len(src), 16)).astype(np.float32) * 1e-3
    rix = np.arange(0, 64, 4)
    a = ora.migrate_shot(grid, vel, vel, src, 30, 2, data, rix, 2, 2)
    b = ora.migrate_shot(grid, vel, vel, src, 30, 2, data, rix, 2, 2, use_reference_kernels=True)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])


# ---------------------------------------------------------------- invariants
def test_stencil_symbol_plane_wave():
    # 4th-order symbol: (2 c2 cos 2kh + 2 c1 cos kh + c0)/h^2, c = (-5/2, 4/3, -1/12)
    nz, nx, h = 16, 64, 5.0
    kx = 2 * np.pi * 3 / (nx * h)
    x = np.arange(nx) * h
    p = np.tile(np.cos(kx * x), (nz, 1))
    vel = np.full((nz, nx), 2000.0)
    dt = 4e-4
    out = np.zeros((nz, nx))
    ora.acoustic_step(p, p, out, vel, None, h, h, dt, False)
    sym = (2 * (-1 / 12) * np.cos(2 * kx * h) + 2 * (4 / 3) * np.cos(kx * h) - 2.5) / h ** 2
    expect = p + (2000.0 ** 2) * dt ** 2 * sym * p  # constant in z: d2z = 0
    np.testing.assert_allclose(out[2:-2, 2:-2], expect[2:-2, 2:-2], rtol=1e-12, atol=1e-12)


def test_constant_field_is_stationary():
    nz, nx = 12, 20
    p = np.full((nz, nx), 0.7, np.float32)
    out = np.zeros_like(p)
    ora.acoustic_step(p, p, out, np.full_like(p, 2500), None, np.float32(5), np.float32(5),
                      np.float32(5e-4), 0)
    np.testing.assert_allclose(out[2:-2, 2:-2], 0.7, rtol=1e-6)


def test_zero_in_zero_out_and_untouched_cells():
    nz, nx = 20, 30
    rng = np.random.default_rng(1)
    z = np.zeros((nz, nx), np.float32)
    out = rng.standard_normal((nz, nx)).astype(np.float32)
    keep = out.copy()
    ora.acoustic_step(z, z, out, np.full_like(z, 2000), ora.build_sponge(nx, nz, 5), 5.0, 5.0,
                      5e-4, 1)
    assert not out[2:-2, 2:-2].any()
    assert not out[0].any()
    np.testing.assert_array_equal(out[1, :2], -keep[2, :2])  # row 1 mirrors stale row-2 edge
    np.testing.assert_array_equal(out[-2:], keep[-2:])
    np.testing.assert_array_equal(out[2:-2, :2], keep[2:-2, :2])


def test_forward_linearity_and_mirror_symmetry():
    nz, nx, dt = 30, 41, 5e-4
    grid = ora.Grid(nx, nz, 5.0, 5.0, dt, 6)
    rng = np.random.default_rng(2)
    half = rng.uniform(1500, 2500, (nz, nx // 2 + 1)).astype(np.float32)
    vel = np.concatenate([half, half[:, -2::-1]], axis=1)
    assert vel.shape == (nz, nx) and np.array_equal(vel, vel[:, ::-1])
    src, _ = ora.ricker(30.0, dt, 0.04)
    rix = np.array([3, 10, 20, 30, 37])
    rec1, s1 = ora.forward(grid, vel, src, 20, 2, rix, 2, 1, True)
    rec2, _ = ora.forward(grid, vel, 2 * src, 20, 2, rix, 2, 1, False)
    np.testing.assert_allclose(rec2, 2 * rec1, rtol=0, atol=1e-30)
    np.testing.assert_array_equal(s1, s1[:, :, ::-1])
    np.testing.assert_array_equal(rec1[:, 0], rec1[:, 4])


def test_fp64_matches_fp32():
    nz, nx, dt = 40, 60, 5e-4
    grid = ora.Grid(nx, nz, 5.0, 5.0, dt, 8)
    rng = np.random.default_rng(4)
    vel = rng.uniform(1500, 2500, (nz, nx)).astype(np.float32)
    src, _ = ora.ricker(25.0, dt, 0.06)
    data = (1e-3 * rng.standard_normal((len(src), 20))).astype(np.float32)
    rix = np.arange(0, 60, 3)
    i32, l32 = ora.migrate_shot(grid, vel, vel, src, 30, 2, data, rix, 2, 2)
    i64, l64 = ora.migrate_shot(grid, vel, vel, src, 30, 2, data, rix, 2, 2, dtype=np.float64)
    assert _rel(i32, i64) < 1e-5
    assert _rel(l32, l64) < 1e-5


def test_ricker_closed_form():
    f0, dt = 2.5, 5e-4
    w, t = ora.ricker(f0, dt, 2.0)
    assert len(w) == 4000 and w.dtype == np.float32
    i0 = int(round(1 / f0 / dt))
    assert abs(w[i0] - 1.0) < 1e-7
    tz = 1 / f0 + 1 / (np.pi * f0 * np.sqrt(2))
    iz = int(np.argmin(np.abs(t - tz)))
    assert abs(w[iz]) < 5e-3
    assert np.argmax(w) == i0


def _ela_case(g, i):
    nx, nz, nsp, fs, nt, k, six, siz, riz = (int(v) for v in g[f"c{i}_ints"])
    dx, dt, alpha = (float(v) for v in g[f"c{i}_flt"])
    grid = ora.Grid(nx, nz, dx, dx, dt, nsp, alpha, bool(fs))
    return grid, nt, k, six, siz, riz


def test_elastic_loops_golden(golden):
    g = golden("elastic_loops")
    for i in range(int(g["ncase"][0])):
        grid, nt, k, six, siz, riz = _ela_case(g, i)
        model = tuple(g[f"c{i}_par"])
        rec, snaps, ill = ora.elastic_forward(grid, model, g[f"c{i}_src"], six, siz,
                                              g[f"c{i}_rix"], riz, k, save=True, illum=True)
        np.testing.assert_array_equal(rec, g[f"c{i}_rec"], err_msg=f"rec {i}")
        np.testing.assert_array_equal(snaps, g[f"c{i}_snaps"], err_msg=f"snaps {i}")
        np.testing.assert_array_equal(ill, g[f"c{i}_illum"], err_msg=f"illum {i}")
        img = np.zeros((grid.nz, grid.nx), np.float32)
        ora.elastic_backward(grid, model, g[f"c{i}_data"], g[f"c{i}_rix"], riz, snaps, img, k)
        np.testing.assert_array_equal(img, g[f"c{i}_image"], err_msg=f"image {i}")


def test_elastic_linearity_and_fp64():
    nz, nx, dt = 30, 40, 4e-4
    grid = ora.Grid(nx, nz, 5.0, 5.0, dt, 6)
    vp = np.full((nz, nx), 2500.0)
    model = ora.elastic_parameters(vp, vp / 1.8, 310.0 * vp ** 0.25)
    src, _ = ora.ricker(25.0, dt, 0.04)
    rix = np.arange(2, nx - 2, 3)
    r1, _, _ = ora.elastic_forward(grid, model, src, 20, 2, rix, 2)
    r2, _, _ = ora.elastic_forward(grid, model, 2 * src, 20, 2, rix, 2)
    np.testing.assert_allclose(r2, 2 * r1, rtol=0, atol=1e-30)
    r64, _, _ = ora.elastic_forward(grid, model, src, 20, 2, rix, 2, dtype=np.float64)
    assert _rel(r1, r64) < 1e-4
    z = np.zeros_like(src)
    r0, _, _ = ora.elastic_forward(grid, model, z, 20, 2, rix, 2)
    assert not r0.any()
