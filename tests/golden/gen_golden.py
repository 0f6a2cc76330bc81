"""Writes the golden vectors in tests/golden/*.npz.

Run in a container that holds the reference tree (it is not needed afterwards):

    make -C oracle            # compiles wave2d.c from source into oracle/_ref/
    python tests/golden/gen_golden.py

Expected outputs come only from
  * the reference C engine, tsw_rtm/c_engine/wave2d.c, compiled from its own
    source with its own compile line (oracle/Makefile), called through ctypes;
  * numpy/scipy evaluating the reference's Python-layer expressions with
    numpy's own dtype rules: the time loops of tsw_rtm/wave_solver.py:146-242
    (injection ``p[iz, ix] += s[it] * dt**2``, receiver recording, snapshot
    list, 3-buffer rotation, snapshot cursor), the illumination of
    tsw_rtm/rtm_engine.py:89 and the post-processing of rtm_engine.py:145-171
    including ``scipy.signal.convolve2d(.., 'same', 'symm')``.
Nothing here calls the oracle restatement, so the fixtures check it.
The elastic step reads one row above the grid with a free surface
(wave2d.c:135,170); it is called on buffers with two zero rows in front.
"""
import ctypes
import os
import sys

import numpy as np
from scipy import signal

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.path.join(ROOT, "oracle", "_ref", "libwave2d_ref.so")

P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float


def load_ref():
    if not os.path.exists(REF):
        sys.exit(f"{REF} missing: run `make -C oracle` where the reference tree is present")
    R = ctypes.CDLL(REF)
    R.acoustic_2d_step.argtypes = [P] * 5 + [I, I, F, F, F, I]
    R.elastic_2d_step.argtypes = [P] * 10 + [I, I, F, F, F, I]
    R.correlate_imaging_condition.argtypes = [P, P, P, I, I, F]
    R.build_cerjan_sponge.argtypes = [P, I, I, I, F, I]
    return R


R = load_ref()


def ptr(a):
    return None if a is None else a.ctypes.data


def ref_sponge(nx, nz, n, alpha, fs):
    sp = np.ones((nz, nx), np.float32)
    R.build_cerjan_sponge(ptr(sp), nx, nz, n, alpha, int(fs))
    return sp


def ref_step(pc, pp, pn, vel, sp, dx, dz, dt, fs):
    nz, nx = pc.shape
    R.acoustic_2d_step(ptr(pc), ptr(pp), ptr(pn), ptr(vel), ptr(sp), nx, nz, dx, dz, dt, int(fs))


class Solver:
    """The state the reference solver holds (wave_solver.py:89-117)."""

    def __init__(self, nx, nz, dx, dz, dt, n_sponge, alpha, fs):
        self.nx, self.nz, self.dx, self.dz, self.dt = nx, nz, float(dx), float(dz), float(dt)
        self.fs = fs
        self.sponge = ref_sponge(nx, nz, n_sponge, alpha, fs)

    def forward(self, vel, src, six, siz, rix, riz, save, k):
        nt = len(src)
        vel = np.ascontiguousarray(vel, np.float32)
        bufs = [np.zeros((self.nz, self.nx), np.float32) for _ in range(3)]
        prv, cur, nxt = bufs
        rec = np.zeros((nt, len(rix)), np.float32)
        snaps = []
        for it in range(nt):
            cur[siz, six] += src[it] * (self.dt ** 2)
            ref_step(cur, prv, nxt, vel, self.sponge, self.dx, self.dz, self.dt, self.fs)
            rec[it, :] = nxt[riz, rix]
            if save and it % k == 0:
                snaps.append(cur.copy())
            prv, cur, nxt = cur, nxt, prv
        return rec, (np.array(snaps) if save else None)

    def backward(self, vel, data, rix, riz, snaps, image, k):
        nt, nrec = data.shape
        vel = np.ascontiguousarray(vel, np.float32)
        prv, cur, nxt = [np.zeros((self.nz, self.nx), np.float32) for _ in range(3)]
        cursor = (nt - 1) // k
        for it in range(nt - 1, -1, -1):
            for r in range(nrec):
                cur[riz, rix[r]] += data[it, r] * (self.dt ** 2)
            ref_step(cur, prv, nxt, vel, self.sponge, self.dx, self.dz, self.dt, self.fs)
            if it % k == 0 and 0 <= cursor < len(snaps):
                s = np.ascontiguousarray(snaps[cursor])
                R.correlate_imaging_condition(ptr(s), ptr(cur), ptr(image), self.nx, self.nz,
                                              self.dt * k)
                cursor -= 1
            prv, cur, nxt = cur, nxt, prv

    def migrate(self, vel_f, vel_b, src, six, siz, data, rix, riz, k):
        _, snaps = self.forward(vel_f, src, six, siz, rix, riz, True, k)
        illum = np.sum(snaps ** 2, axis=0) * (self.dt * k)
        img = np.zeros((self.nz, self.nx), np.float32)
        self.backward(vel_b, data, rix, riz, snaps, img, k)
        return img, illum, snaps


def laplacian(image, dx, dz):
    kern = np.array([[0.0, 1.0 / (dz * dz), 0.0],
                     [1.0 / (dx * dx), -2.0 / (dx * dx) - 2.0 / (dz * dz), 1.0 / (dx * dx)],
                     [0.0, 1.0 / (dz * dz), 0.0]], dtype=np.float32)
    return signal.convolve2d(image, kern, mode="same", boundary="symm")


def postprocess(sponge, stacked, illum, dx, dz, mute):
    stacked = stacked.copy()
    illum = illum.copy()
    mask = sponge < 0.99
    if mute:
        stacked[mask] = 0.0
        illum[mask] = 0.0
    eps = 1e-4 * np.max(illum) if np.max(illum) > 0 else 1e-12
    norm = stacked / (np.sqrt(illum) + eps)
    lap = laplacian(stacked, dx, dz)
    if mute:
        lap[mask] = 0.0
    return stacked, lap, norm


def smooth_model(rng, nz, nx, lo=1500.0, hi=2500.0):
    z = np.arange(nz)[:, None] / nz
    x = np.arange(nx)[None, :] / nx
    v = lo + (hi - lo) * (z + 0.2 * x) + 30.0 * rng.standard_normal((nz, nx))
    return np.clip(v, lo, hi + 100).astype(np.float32)


def gen_sponge():
    out = {}
    cases = [(40, 24, 5, 0.0015, 1), (40, 24, 5, 0.0015, 0), (53, 37, 9, 0.01, 1),
             (400, 200, 30, 0.0015, 1), (64, 48, 25, 0.0015, 0), (17, 13, 3, 0.2, 1)]
    for i, (nx, nz, n, a, fs) in enumerate(cases):
        out[f"c{i}_args"] = np.array([nx, nz, n, fs], np.int64)
        out[f"c{i}_alpha"] = np.array([a], np.float32)
        out[f"c{i}_out"] = ref_sponge(nx, nz, n, a, fs)
    out["ncase"] = np.array([len(cases)])
    np.savez_compressed(os.path.join(HERE, "sponge.npz"), **out)


def gen_acoustic_step(rng):
    out = {}
    cases = [(40, 24, 5.0, 5.0, 5e-4, 1, True), (40, 24, 5.0, 5.0, 5e-4, 0, True),
             (53, 37, 7.0, 4.0, 6e-4, 1, False), (53, 37, 7.0, 4.0, 6e-4, 0, False),
             (5, 5, 5.0, 5.0, 5e-4, 1, True), (130, 70, 20.0, 20.0, 2e-3, 1, True)]
    # grids with no interior column (nx < 5): only the free-surface rows are
    # written; drawn from a generator of their own so the other fixtures keep
    # their inputs
    narrow = [(3, 6, 5.0, 5.0, 5e-4, 1, True), (4, 3, 5.0, 5.0, 5e-4, 1, False),
              (1, 5, 5.0, 5.0, 5e-4, 1, True), (4, 7, 5.0, 5.0, 5e-4, 0, True)]
    rng_narrow = np.random.default_rng(20251205)
    cases += narrow
    for i, (nx, nz, dx, dz, dt, fs, use_sp) in enumerate(cases):
        if i == len(cases) - len(narrow):
            rng = rng_narrow
        pc = rng.standard_normal((nz, nx)).astype(np.float32)
        pp = rng.standard_normal((nz, nx)).astype(np.float32)
        pn = rng.standard_normal((nz, nx)).astype(np.float32)  # untouched cells must survive
        vel = rng.uniform(1500, 2600, (nz, nx)).astype(np.float32)
        sp = ref_sponge(nx, nz, 4, 0.05, fs) if use_sp else None
        out[f"c{i}_in"] = np.stack([pc, pp, pn, vel])
        out[f"c{i}_sponge"] = sp if sp is not None else np.zeros((0,), np.float32)
        out[f"c{i}_par"] = np.array([dx, dz, dt], np.float32)
        out[f"c{i}_fs"] = np.array([fs])
        res = pn.copy()
        ref_step(pc, pp, res, vel, sp, dx, dz, dt, fs)
        out[f"c{i}_out"] = res
    out["ncase"] = np.array([len(cases)])
    np.savez_compressed(os.path.join(HERE, "acoustic_step.npz"), **out)


def gen_elastic_step(rng):
    out = {}
    cases = [(40, 24, 5.0, 5.0, 4e-4, 1, True, 1), (40, 24, 5.0, 5.0, 4e-4, 0, True, 1),
             (33, 21, 6.0, 4.0, 3e-4, 1, False, 1), (48, 32, 5.0, 5.0, 4e-4, 1, True, 6),
             (48, 32, 5.0, 5.0, 4e-4, 0, False, 6)]
    for i, (nx, nz, dx, dz, dt, fs, use_sp, nsteps) in enumerate(cases):
        vp = rng.uniform(1800, 3500, (nz, nx)).astype(np.float32)
        vs = (vp / 1.8).astype(np.float32)
        rho = (310.0 * vp ** 0.25).astype(np.float32)
        mu = (rho * vs * vs).astype(np.float32)
        lam = (rho * vp * vp - 2 * mu).astype(np.float32)
        bx = (1.0 / rho).astype(np.float32)
        bz = (1.0 / rho).astype(np.float32)
        sp = ref_sponge(nx, nz, 4, 0.05, fs) if use_sp else None
        # two zero rows in front of each state field (row -1 reads 0)
        pad = [np.zeros((nz + 2, nx), np.float32) for _ in range(5)]
        for f in pad:
            f[2:] = (1e-3 * rng.standard_normal((nz, nx))).astype(np.float32)
        init = np.stack([f[2:].copy() for f in pad])
        for _ in range(nsteps):
            R.elastic_2d_step(*[f.ctypes.data + 2 * nx * 4 for f in pad], ptr(lam), ptr(mu),
                              ptr(bx), ptr(bz), ptr(sp), nx, nz, dx, dz, dt, fs)
        assert all(not f[:2].any() for f in pad)
        out[f"c{i}_init"] = init
        out[f"c{i}_par_fields"] = np.stack([lam, mu, bx, bz])
        out[f"c{i}_sponge"] = sp if sp is not None else np.zeros((0,), np.float32)
        out[f"c{i}_par"] = np.array([dx, dz, dt], np.float32)
        out[f"c{i}_ints"] = np.array([fs, nsteps])
        out[f"c{i}_out"] = np.stack([f[2:] for f in pad])
    out["ncase"] = np.array([len(cases)])
    np.savez_compressed(os.path.join(HERE, "elastic_step.npz"), **out)


def gen_correlate(rng):
    out = {}
    for i, (nx, nz, w) in enumerate([(40, 24, 1e-3), (7, 3, 0.37)]):
        a = rng.standard_normal((nz, nx)).astype(np.float32)
        b = rng.standard_normal((nz, nx)).astype(np.float32)
        img = rng.standard_normal((nz, nx)).astype(np.float32)
        out[f"c{i}_in"] = np.stack([a, b, img])
        out[f"c{i}_w"] = np.array([w], np.float32)
        R.correlate_imaging_condition(ptr(a), ptr(b), ptr(img), nx, nz, w)
        out[f"c{i}_out"] = img
    out["ncase"] = np.array([2])
    np.savez_compressed(os.path.join(HERE, "correlate.npz"), **out)


def gen_loops(rng):
    """Forward records/snapshots and backward images on tiny grids."""
    out = {}
    cases = [
        # nx, nz, dx,  dt,   nsp, fs, nt, k, six, siz, rix,                       riz
        (40, 24, 5.0, 5e-4, 5, 1, 60, 1, 20, 2, [0, 1, 5, 9, 9, 20, 38, 39], 2),
        (40, 24, 5.0, 5e-4, 5, 1, 61, 2, 11, 3, list(range(0, 40, 3)), 3),
        (33, 20, 4.0, 4e-4, 4, 0, 45, 3, 16, 2, [2, 4, 4, 4, 30], 2),
        (16, 12, 5.0, 5e-4, 2, 1, 20, 2, 0, 0, [0, 15], 1),
    ]
    for i, (nx, nz, dx, dt, nsp, fs, nt, k, six, siz, rix, riz) in enumerate(cases):
        s = Solver(nx, nz, dx, dx, dt, nsp, 0.0015 * 20, fs)
        vel_t = smooth_model(rng, nz, nx)
        vel_m = smooth_model(rng, nz, nx)
        t = np.arange(nt) * dt
        f0 = 25.0
        tau = t - 1.0 / f0
        src = ((1 - 2 * (np.pi * f0 * tau) ** 2) * np.exp(-(np.pi * f0 * tau) ** 2)).astype(np.float32)
        rix = np.array(rix)
        rec, snaps = s.forward(vel_t, src, six, siz, rix, riz, True, k)
        data = (rec + 0.01 * rng.standard_normal(rec.shape)).astype(np.float32)
        img = (0.1 * rng.standard_normal((nz, nx))).astype(np.float32)  # accumulate in place
        img0 = img.copy()
        s.backward(vel_m, data, rix, riz, snaps, img, k)
        out[f"c{i}_ints"] = np.array([nx, nz, nsp, fs, nt, k, six, siz, riz])
        out[f"c{i}_flt"] = np.array([dx, dt, 0.0015 * 20])
        out[f"c{i}_rix"] = rix.astype(np.int32)
        out[f"c{i}_vel"] = np.stack([vel_t, vel_m])
        out[f"c{i}_src"] = src
        out[f"c{i}_rec"] = rec
        out[f"c{i}_snaps"] = snaps
        out[f"c{i}_data"] = data
        out[f"c{i}_img0"] = img0
        out[f"c{i}_img"] = img
    # snapshot list shorter than the cursor start: the reference never images
    s = Solver(16, 12, 5.0, 5.0, 5e-4, 2, 0.03, 1)
    img = np.zeros((12, 16), np.float32)
    data = rng.standard_normal((20, 2)).astype(np.float32)
    snaps = rng.standard_normal((3, 12, 16)).astype(np.float32)
    s.backward(smooth_model(rng, 12, 16), data, np.array([3, 7]), 2, snaps, img, 2)
    out["short_img"] = img
    out["ncase"] = np.array([len(cases)])
    np.savez_compressed(os.path.join(HERE, "loops.npz"), **out)


def gen_migrate(rng):
    """migrate_shot and migrate_multiple_shots post-processing, 3 shots."""
    nx, nz, dx, dz, dt, nsp, nt, k = 48, 30, 5.0, 5.0, 5e-4, 6, 80, 2
    s = Solver(nx, nz, dx, dz, dt, nsp, 0.0015, 1)
    vel_t = smooth_model(rng, nz, nx)
    vel_f = smooth_model(rng, nz, nx)
    vel_b = smooth_model(rng, nz, nx)
    rix = np.arange(0, nx, 2)
    t = np.arange(nt) * dt
    f0 = 30.0
    tau = t - 1.0 / f0
    src = ((1 - 2 * (np.pi * f0 * tau) ** 2) * np.exp(-(np.pi * f0 * tau) ** 2)).astype(np.float32)
    sixs = [10, 24, 37]
    stacked = np.zeros((nz, nx), np.float32)
    illum = np.zeros((nz, nx), np.float32)
    out = {}
    for j, six in enumerate(sixs):
        rec, _ = s.forward(vel_t, src, six, 2, rix, 2, False, 1)
        img, ill, _ = s.migrate(vel_f, vel_b, src, six, 2, rec, rix, 2, k)
        stacked += img
        illum += ill
        out[f"s{j}_data"] = rec
        out[f"s{j}_img"] = img
        out[f"s{j}_ill"] = ill
    raw, lap, norm = postprocess(s.sponge, stacked, illum, dx, dz, True)
    raw_nm, lap_nm, norm_nm = postprocess(s.sponge, stacked, illum, dx, dz, False)
    out.update(ints=np.array([nx, nz, nsp, nt, k, 2, 2]), flt=np.array([dx, dz, dt, 0.0015]),
               sixs=np.array(sixs), rix=rix.astype(np.int32), vel=np.stack([vel_t, vel_f, vel_b]),
               src=src, raw=raw, lap=lap, norm=norm, raw_nm=raw_nm, lap_nm=lap_nm,
               norm_nm=norm_nm, stacked=stacked, illum=illum)
    # Laplacian of random images (scipy, float32 kernel), odd shapes included
    for i, (a, b) in enumerate([(30, 48), (7, 5), (1, 9), (64, 33)]):
        im = rng.standard_normal((a, b)).astype(np.float32)
        out[f"lap{i}_in"] = im
        out[f"lap{i}_out"] = laplacian(im, 5.0 + i, 4.0 + 2 * i)
        out[f"lap{i}_d"] = np.array([5.0 + i, 4.0 + 2 * i])
    np.savez_compressed(os.path.join(HERE, "migrate.npz"), **out)


def gen_elastic_loops(rng):
    """Elastic driver loops (build-defined conventions: vertical force
    vz += (dt*bz)*s, vertical strain-rate receivers, vz imaging) around the
    reference's elastic_2d_step compiled from source, on padded buffers."""
    f32 = np.float32
    c1, c2 = f32(9) / f32(8), f32(-1) / f32(24)
    out = {}
    cases = [(40, 30, 5.0, 4e-4, 5, 1, 50, 2, 20, 2, [0, 3, 3, 10, 20, 39], 2),
             (33, 26, 4.0, 3e-4, 4, 0, 41, 3, 5, 3, [2, 7, 30], 4),
             (48, 36, 5.0, 4e-4, 6, 1, 60, 1, 24, 0, [1, 24, 46], 1)]
    for i, (nx, nz, dx, dt, nsp, fs, nt, k, six, siz, rix, riz) in enumerate(cases):
        sponge = ref_sponge(nx, nz, nsp, 0.02, fs)
        vp = (2000 + 800 * rng.random((nz, nx))).astype(f32)
        rho = (310.0 * vp.astype(np.float64) ** 0.25)
        vs = vp / 1.8
        mu = (rho * vs * vs).astype(f32)
        lam = (rho * vp * vp - 2 * rho * vs * vs).astype(f32)
        b = (1.0 / rho).astype(f32)
        par = [np.ascontiguousarray(a) for a in (lam, mu, b, b)]
        src = (rng.standard_normal(nt) * 1e3).astype(f32)
        rix = np.array(rix)
        dtf, rdz = f32(dt), f32(1) / f32(dx)

        def run(inject, it_order, side, record):
            pad = [np.zeros((nz + 2, nx), f32) for _ in range(5)]
            vz = pad[1][2:]
            for it in it_order:
                inject(vz, it)
                side(vz, it)
                R.elastic_2d_step(*[p.ctypes.data + 2 * nx * 4 for p in pad],
                                  *[ptr(a) for a in par], ptr(sponge), nx, nz, dx, dx, dt, fs)
                record(vz, it)

        def strain(vz, iz, ix):
            def at(z):
                return vz[z, ix] if 0 <= z < nz else f32(0)
            return (c1 * (at(iz) - at(iz - 1)) + c2 * (at(iz + 1) - at(iz - 2))) * rdz

        rec = np.zeros((nt, len(rix)), f32)
        snaps = []
        illum = np.zeros((nz, nx), f32)

        def inj_src(vz, it):
            vz[siz, six] = vz[siz, six] + (dtf * par[3][siz, six]) * src[it]

        def side_fwd(vz, it):
            if it % k == 0:
                snaps.append(vz.copy())
                illum[...] = illum + vz * vz

        def rec_fwd(vz, it):
            for r in range(len(rix)):
                rec[it, r] = strain(vz, riz, rix[r])

        run(inj_src, range(nt), side_fwd, rec_fwd)
        snaps = np.array(snaps)
        data = (rec + f32(1e-9) * rng.standard_normal(rec.shape).astype(f32)).astype(f32)
        image = np.zeros((nz, nx), f32)
        w = f32(dt * k)
        cursor = [(nt - 1) // k]

        def inj_data(vz, it):
            for r in range(len(rix)):
                vz[riz, rix[r]] = vz[riz, rix[r]] + (dtf * par[3][riz, rix[r]]) * data[it, r]

        def side_bwd(vz, it):
            if it % k == 0 and 0 <= cursor[0] < len(snaps):
                image[...] = image + (snaps[cursor[0]] * vz) * w
                cursor[0] -= 1

        run(inj_data, range(nt - 1, -1, -1), side_bwd, lambda vz, it: None)
        out[f"c{i}_ints"] = np.array([nx, nz, nsp, fs, nt, k, six, siz, riz])
        out[f"c{i}_flt"] = np.array([dx, dt, 0.02])
        out[f"c{i}_par"] = np.stack(par)
        out[f"c{i}_src"] = src
        out[f"c{i}_rix"] = rix.astype(np.int32)
        out[f"c{i}_rec"] = rec
        out[f"c{i}_snaps"] = snaps
        out[f"c{i}_illum"] = illum
        out[f"c{i}_data"] = data
        out[f"c{i}_image"] = image
    out["ncase"] = np.array([len(cases)])
    np.savez_compressed(os.path.join(HERE, "elastic_loops.npz"), **out)


def main():
    rng = np.random.default_rng(20251203)
    gen_sponge()
    gen_acoustic_step(rng)
    gen_elastic_step(rng)
    gen_correlate(rng)
    gen_loops(rng)
    gen_migrate(rng)
    gen_elastic_loops(np.random.default_rng(20251204))
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
