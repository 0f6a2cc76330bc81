"""Serving loop: each request's result equals migrating it alone (bit for
bit), whether it runs alone or in a mix of uneven acoustic and elastic requests
that queue against a tight memory budget; admission respects the slot and
budget limits; cancellation and an injected admission failure affect only the
request concerned."""
import time

import numpy as np
import pytest

from das_amd import _native as N

pytestmark = pytest.mark.gpu

DX, DT, NSP, K = 5.0, 4e-4, 8, 2


def _requests(seed=3):
    from das_amd.wave_solver import WaveSolver2D, ricker_wavelet
    from das_amd.workloads import (ElasticShot, Shot, correlogram_gather, elastic_from_vp,
                                   faulted_two_layer, layered_faulted_vp, smooth)

    rng = np.random.default_rng(seed)
    reqs = []
    specs = [("a", 64, 200, 130), ("e", 40, 96, 70), ("a", 96, 160, 90), ("a", 48, 256, 151),
             ("e", 56, 120, 61), ("a", 70, 300, 77)]
    for kind, nz, nx, nt in specs:
        six = int(rng.integers(5, nx - 5))
        rix = np.arange(1, nx - 1, 2).astype(np.int32)
        src, _ = ricker_wavelet(25.0, DT, nt * DT)
        if kind == "a":
            v = smooth(faulted_two_layer(nz, nx, DX, seed=int(rng.integers(99))), 30.0, DX)
            data = correlogram_gather(rng, nt, DT, (rix - six) * DX, f0=25.0, f_lo=5, f_hi=50)
            reqs.append(Shot(nx=nx, nz=nz, nt=nt, vel_fwd=v, vel_bwd=v, src=src, six=six,
                             siz=2, data=data, rix=rix, riz=2))
        else:
            vp = layered_faulted_vp(nz, nx, DX, seed=int(rng.integers(99)))
            mt, mm = elastic_from_vp(vp), elastic_from_vp(smooth(vp, 30.0, DX))
            solver = WaveSolver2D(nx, nz, DX, DX, DT, n_sponge=NSP, solver_type="elastic")
            data, _ = solver.forward_propagate_elastic(mt, src, six, 2, rix, 2)
            reqs.append(ElasticShot(nx=nx, nz=nz, nt=nt, model_fwd=mm, model_bwd=mm, src=src,
                                    six=six, siz=2, data=data, rix=rix, riz=2))
    return reqs


def _alone(s):
    from das_amd.rtm_engine import TSWRTMEngine
    from das_amd.velocity_flood import FloodModels
    from das_amd.wave_solver import WaveSolver2D

    solver = WaveSolver2D(s.nx, s.nz, DX, DX, DT, n_sponge=NSP)
    eng = TSWRTMEngine(solver)
    if getattr(s, "kind", "acoustic") == "elastic":
        return eng.migrate_shot_elastic(s.model_fwd, s.model_bwd, s.src, s.six, s.siz, s.data,
                                        s.rix, s.riz, K)
    return eng.migrate_shot(FloodModels(s.vel_fwd, s.vel_fwd, s.vel_bwd, 0, 0), s.src, s.six,
                            s.siz, s.data, s.rix, s.riz, K)


@pytest.fixture(scope="module")
def reqs():
    N.check(N.lib().tsw_init(0))
    r = _requests()
    return r, [_alone(s) for s in r]


def _server(**kw):
    from das_amd.serve import RTMServer

    args = dict(n_sponge=NSP, k=K, budget_bytes=4e9, max_active=8, kernel=11, tick_steps=40)
    args.update(kw)
    return RTMServer(DX, DX, DT, **args)


def test_each_request_alone_equals_single_shot(reqs):
    shots, expect = reqs
    srv = _server()
    for s, (img, ill) in zip(shots, expect):
        rid = srv.submit(s)
        assert srv.wait(rid)["state"] == "done"
        got = srv.result(rid)
        np.testing.assert_array_equal(got[0], img)
        np.testing.assert_array_equal(got[1], ill)
    srv.close()


def test_mix_under_tight_budget_equals_alone(reqs):
    from das_amd.autotune import free_device_bytes  # noqa: F401  (device query works)

    shots, expect = reqs
    need = []
    for s in shots:
        n = s.nx * s.nz
        per = (10 if getattr(s, "kind", "") != "elastic" else 20) * 4 * n
        need.append(per + 4 * ((s.nt - 1) // K + 1) * n)
    budget = int(2.2 * max(need))  # forces queueing: at most a few requests fit at once
    srv = _server(budget_bytes=budget, max_active=3, tick_steps=17)
    ids = []
    for i, s in enumerate(shots):
        ids.append(srv.submit(s))
        time.sleep(0.002 * (i % 3))  # staggered arrivals while others run
    polls = [srv.wait(r) for r in ids]
    assert all(p["state"] == "done" for p in polls)
    for rid, (img, ill) in zip(ids, expect):
        got = srv.result(rid)
        np.testing.assert_array_equal(got[0], img)
        np.testing.assert_array_equal(got[1], ill)
    # admission order = arrival order, and never more than 3 running at once
    admits = sorted(p["t_admit"] for p in polls)
    assert [p["t_admit"] for p in polls] == admits
    events = sorted([(p["t_admit"], 1) for p in polls] + [(p["t_done"], -1) for p in polls])
    run = peak = 0
    for _, d in events:
        run += d
        peak = max(peak, run)
    assert peak <= 3
    st = srv.stats()
    assert st["completed"] == len(shots) and st["failed"] == 0 and st["mem_used"] == 0
    srv.close()


def test_budget_rejection_cancel_and_injected_failure(reqs):
    shots, expect = reqs
    srv = _server(budget_bytes=1e6)
    with pytest.raises(MemoryError):
        srv.submit(shots[0])
    srv.close()
    srv = _server(max_active=1)
    a = srv.submit(shots[3])
    b = srv.submit(shots[0])
    srv.cancel(b)  # still queued behind a
    assert srv.wait(a)["state"] == "done"
    assert srv.poll(b)["state"] == "cancelled"
    np.testing.assert_array_equal(srv.result(a)[0], expect[3][0])
    N.lib().tsw_debug_fail_next(1)
    c = srv.submit(shots[2])
    d = srv.submit(shots[1])
    assert srv.wait(c)["state"] == "failed"
    assert srv.wait(d)["state"] == "done"
    got = srv.result(d)
    np.testing.assert_array_equal(got[0], expect[1][0])
    np.testing.assert_array_equal(got[1], expect[1][1])
    with pytest.raises(N.TswError):
        srv.result(c)
    # reported once, then forgotten: failed and cancelled records do not pile up
    with pytest.raises(N.TswError, match="cancelled"):
        srv.result(b)
    for rid in (b, c):
        with pytest.raises(ValueError, match="unknown request"):
            srv.poll(rid)
    st = srv.stats()
    assert st["failed"] == 1 and st["cancelled"] == 1 and st["mem_used"] == 0
    srv.close()


def test_lone_failed_admission_wakes_its_waiter(reqs):
    """A request whose admission fails while nothing else runs (so no tick
    follows) is reported failed at once, not when the wait times out."""
    shots, _ = reqs
    srv = _server(max_active=1)
    N.lib().tsw_debug_fail_next(1)
    rid = srv.submit(shots[2])
    t0 = time.monotonic()
    assert srv.wait(rid, timeout=60.0)["state"] == "failed"
    assert time.monotonic() - t0 < 10.0
    srv.close()


def test_resident_model_windows_equal_uploaded_models():
    """Requests that name a window of the server's resident models give the
    results of the same requests carrying copies of those windows."""
    from dataclasses import replace

    from das_amd.wave_solver import ricker_wavelet
    from das_amd.workloads import (ElasticShot, Shot, correlogram_gather, elastic_from_vp,
                                   layered_faulted_vp, smooth)

    rng = np.random.default_rng(21)
    gnz, gnx = 90, 420
    vp = layered_faulted_vp(gnz, gnx, DX, seed=3)
    vel = smooth(vp, 30.0, DX)
    ela = elastic_from_vp(vel)
    reqs = []
    for kind, nz, nx, nt, x0, z0 in [("a", 64, 200, 120, 0, 0), ("e", 48, 96, 60, 310, 20),
                                     ("a", 90, 160, 80, 260, 0), ("e", 40, 120, 70, 7, 50),
                                     ("a", 50, 256, 90, 164, 33)]:
        six = int(rng.integers(5, nx - 5))
        rix = np.arange(1, nx - 1, 3).astype(np.int32)
        src, _ = ricker_wavelet(25.0, DT, nt * DT)
        data = correlogram_gather(rng, nt, DT, (rix - six) * DX, f0=25.0, f_lo=5, f_hi=50)
        win = (slice(z0, z0 + nz), slice(x0, x0 + nx))
        if kind == "a":
            v = np.ascontiguousarray(vel[win])
            reqs.append(Shot(nx=nx, nz=nz, nt=nt, vel_fwd=v, vel_bwd=v, src=src, six=six,
                             siz=2, data=data, rix=rix, riz=2, x0=x0, z0=z0))
        else:
            m = tuple(np.ascontiguousarray(a[win]) for a in ela)
            reqs.append(ElasticShot(nx=nx, nz=nz, nt=nt, model_fwd=m, model_bwd=m, src=src,
                                    six=six, siz=2, data=data, rix=rix, riz=2, x0=x0, z0=z0))
    srv = _server(max_active=3, tick_steps=23)
    with pytest.raises(ValueError, match="no resident"):
        srv.submit(replace(reqs[0], vel_fwd=None, vel_bwd=None))
    srv.set_models(vel, vel, ela, ela)
    with pytest.raises(ValueError, match="outside"):
        srv.submit(replace(reqs[0], vel_fwd=None, vel_bwd=None, x0=gnx - 100))
    up = [srv.submit(s) for s in reqs]
    res = [srv.submit(replace(s, vel_fwd=None, vel_bwd=None) if isinstance(s, Shot)
                      else replace(s, model_fwd=None, model_bwd=None)) for s in reqs]
    for a, b in zip(up, res):
        assert srv.wait(a)["state"] == "done" and srv.wait(b)["state"] == "done"
        ia, la = srv.result(a)
        ib, lb = srv.result(b)
        np.testing.assert_array_equal(ia, ib)
        np.testing.assert_array_equal(la, lb)
        assert np.abs(ia).max() > 0
    srv.close()
