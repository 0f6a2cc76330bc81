"""Serving virtual-shot migration requests on the HIP engine.

``RTMServer`` wraps the C++ serving loop (include/tswrtm.h, tsw_server_*):
requests of uneven grid size and length (acoustic or elastic shots) are
admitted against a device-memory budget as slots free up, advanced together by
continuous batching, and timed on a monotonic clock.  ``run_poisson`` drives a
server with a seeded Poisson arrival stream and reports latency percentiles and
throughput.
"""
from __future__ import annotations

import ctypes
import time
from typing import Sequence

import numpy as np

from . import _native as N
from .batch import fill_desc

STATES = ("queued", "running", "done", "failed", "cancelled")


class RTMServer:
    def __init__(self, dx: float, dz: float, dt: float, n_sponge: int = 30,
                 sponge_alpha: float = 0.0015, free_surface_top: bool = True, k: int = 2,
                 budget_bytes: float = 200e9, max_active: int = 32, kernel: int = 15,
                 tick_steps: int = 100):
        h = N.P()
        N.check(N.lib().tsw_server_create(float(dx), float(dz), float(dt), int(n_sponge),
                                          float(np.float32(sponge_alpha)),
                                          1 if free_surface_top else 0, int(k),
                                          int(budget_bytes), int(max_active), int(kernel),
                                          int(tick_steps), ctypes.byref(h)))
        self._h = h
        self._shape = {}

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib().tsw_server_destroy(self._h)
            self._h = None

    __del__ = close

    def set_models(self, vel_fwd=None, vel_bwd=None, ela_fwd=None, ela_bwd=None):
        """Uploads models shared by requests (each (gnz, gnx)); a request
        whose model fields are None migrates in its window (z0, x0) of them.
        Only while no request is queued or running."""
        arrs = [a for a in (vel_fwd, vel_bwd) if a is not None]
        arrs += [a for m in (ela_fwd, ela_bwd) if m is not None for a in m]
        if not arrs:
            raise ValueError("no models given")
        shape = np.asarray(arrs[0]).shape
        keep = [N.f32(a) for a in arrs]
        if any(a.shape != shape or a.ndim != 2 for a in keep):
            raise ValueError("resident models must all have the same (nz, nx) shape")
        vf = N.f32(vel_fwd) if vel_fwd is not None else None
        vb = N.f32(vel_bwd) if vel_bwd is not None else None
        ef = [N.f32(a) for a in ela_fwd] if ela_fwd is not None else None
        eb = [N.f32(a) for a in ela_bwd] if ela_bwd is not None else None
        PF = (N.P * 4)(*[N.ptr(a) for a in ef]) if ef is not None else None
        PB = (N.P * 4)(*[N.ptr(a) for a in eb]) if eb is not None else None
        N.check(N.lib().tsw_server_set_models(self._h, shape[1], shape[0], N.ptr(vf), N.ptr(vb),
                                              PF, PB))

    def submit(self, shot) -> int:
        keep = []
        d = N.ShotDesc()
        fill_desc(d, shot, keep)
        rid = N.LL()
        N.check(N.lib().tsw_server_submit(self._h, ctypes.byref(d), ctypes.byref(rid)))
        self._shape[rid.value] = (shot.nz, shot.nx)
        return rid.value

    def poll(self, rid: int) -> dict:
        st = N.I()
        ts, ta, td = N.D(), N.D(), N.D()
        N.check(N.lib().tsw_server_poll(self._h, int(rid), ctypes.byref(st), ctypes.byref(ts),
                                        ctypes.byref(ta), ctypes.byref(td)))
        return {"state": STATES[st.value], "t_submit": ts.value, "t_admit": ta.value,
                "t_done": td.value}

    def wait(self, rid: int, timeout: float = 600.0) -> dict:
        N.check(N.lib().tsw_server_wait(self._h, int(rid), float(timeout)))
        return self.poll(rid)

    def result(self, rid: int):
        nz, nx = self._shape.pop(rid)
        img = np.empty((nz, nx), np.float32)
        ill = np.empty_like(img)
        N.check(N.lib().tsw_server_result(self._h, int(rid), N.ptr(img), N.ptr(ill)))
        return img, ill

    def cancel(self, rid: int):
        N.check(N.lib().tsw_server_cancel(self._h, int(rid)))

    def stats(self) -> dict:
        v = [N.LL() for _ in range(7)]
        busy = N.D()
        N.check(N.lib().tsw_server_stats(self._h, *[ctypes.byref(x) for x in v],
                                         ctypes.byref(busy)))
        keys = ("completed", "failed", "cancelled", "queued", "running", "mem_used",
                "cell_updates")
        out = {k: x.value for k, x in zip(keys, v)}
        out["busy_s"] = busy.value
        return out


def run_poisson(server: RTMServer, shots: Sequence, rate_hz: float, seed: int = 0,
                keep_results: bool = False, log=None, make=None) -> dict:
    """Submits `shots` at seeded exponential inter-arrival times (mean
    1/rate_hz) on the host clock, waits for all, and reports latency
    (arrival -> image ready, server clock) percentiles and throughput.
    With `make`, `shots` are lightweight specs and make(spec) builds each
    request just before its arrival (the server copies it on submission)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_hz, len(shots))
    arrivals = np.cumsum(gaps) - gaps[0]
    t0 = time.perf_counter()
    ids = []
    for i, (s, ta) in enumerate(zip(shots, arrivals)):
        delay = ta - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ids.append(server.submit(make(s) if make else s))
        if log and (i + 1) % 50 == 0:
            log(f"served: submitted {i + 1}/{len(shots)}")
    lat, results, failed = [], {}, 0
    t_first = None
    t_last = 0.0
    for rid in ids:
        p = server.wait(rid)
        if p["state"] != "done":
            failed += 1
            continue
        lat.append(p["t_done"] - p["t_submit"])
        t_first = p["t_submit"] if t_first is None else min(t_first, p["t_submit"])
        t_last = max(t_last, p["t_done"])
        img = server.result(rid)
        if keep_results:
            results[rid] = img
    wall = time.perf_counter() - t0
    lat = np.array(lat)
    span = t_last - (t_first or 0.0)
    cu = sum(2 * s.nt * s.nx * s.nz for s in shots)
    st = server.stats()
    out = {"requests": len(shots), "failed": failed, "rate_hz": rate_hz,
           "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
           "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
           "mean_latency_s": float(lat.mean()) if len(lat) else None,
           "shots_per_s": len(lat) / span if span > 0 else None,
           "gcell_updates_per_s": cu / span / 1e9 if span > 0 else None,
           "gpu_busy_s": st["busy_s"], "wall_s": wall, "span_s": span,
           "ids": ids}
    if keep_results:
        out["results"] = results
    return out
