"""Autotuner for the batched RTM kernels.

Chooses the step kernel (the per-step kernel 0 or one of the temporally
blocked configurations: steps per launch T and tile shape) and the number of
concurrent slots for a batch of shots, by timing candidates on a shortened
copy of the batch (same grids and models, nt scaled down).  Results are cached
in das_amd/tuning_cache.json keyed by device and workload signature, so a
tuned workload is not re-measured.
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from dataclasses import replace

import numpy as np

from . import _native as N

CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_cache.json")
DEFAULT_KERNELS = (0, 7, 9, 10, 11, 12, 15, 16, 25, 26, 27, 28)


def device_name() -> str:
    buf = ctypes.create_string_buffer(128)
    N.check(N.lib().tsw_device_name(buf, 128))
    return buf.value.decode()


def free_device_bytes() -> int:
    f, t = N.LL(), N.LL()
    N.check(N.lib().tsw_mem_info(ctypes.byref(f), ctypes.byref(t)))
    return f.value


def workload_key(shots, k) -> str:
    nz = sorted({s.nz for s in shots})
    nx = [s.nx for s in shots]
    nt = [s.nt for s in shots]
    return (f"{device_name()}|n={len(shots)}|nz={nz}|nx={min(nx)}-{max(nx)}|"
            f"nt={min(nt)}-{max(nt)}|k={k}")


def state_bytes(shots, k, slots) -> int:
    """Device bytes a batch needs: per-slot wavefields and snapshot arena
    (sized by the slot's largest shot, longest-first packing approximated by
    the `slots` largest shots), plus per-shot models, data and images."""
    per_shot = sum(4 * (4 * s.nx * s.nz + s.nt * (len(s.rix) + 1)) for s in shots)
    big = sorted((((s.nt - 1) // k + 1 + 6) * s.nx * s.nz * 4 for s in shots), reverse=True)
    return per_shot + sum(big[:max(1, min(slots, len(shots)))])


def _shorten(s, nt):
    return replace(s, nt=nt, src=np.ascontiguousarray(s.src[:nt]),
                   data=np.ascontiguousarray(s.data[:nt]))


def load_cache() -> dict:
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def tune(shots, dx, dt, k=2, n_sponge=30, kernels=DEFAULT_KERNELS, slot_options=(16, 32, 64),
         nt_scale=0.12, memory_fraction=0.9, use_cache=True, write_cache=True, log=None,
         graph_steps=20000):
    """Returns {"kernel", "slots", "measured": {...}} for this batch of shots,
    each candidate timed as the batch runs it (hipGraphs of `graph_steps`
    time steps)."""
    from .batch import ShotBatch

    key = workload_key(shots, k)
    cache = load_cache()
    if use_cache and key in cache:
        return cache[key]
    budget = memory_fraction * free_device_bytes()
    short = [_shorten(s, max(24, int(s.nt * nt_scale))) for s in shots]
    scale = sum(2 * s.nt * s.nx * s.nz for s in short)
    results = {}
    for slots in slot_options:
        slots = min(slots, len(shots))
        if state_bytes(shots, k, slots) > budget:
            continue
        # two passes, the second in reverse order, each candidate's best run
        # kept: a candidate's rate depends by a few percent on what ran
        # before it
        for order in (list(kernels), list(kernels)[::-1]):
            for kern in order:
                b = ShotBatch(short, dx, dx, dt, n_sponge=n_sponge, k=k, slots=slots,
                              kernel=kern, graph_steps=graph_steps)
                b.run("batched")
                el = float("inf")
                for _ in range(2):
                    t0 = time.perf_counter()
                    b.run("batched")
                    el = min(el, time.perf_counter() - t0)
                b.close()
                key_ks = f"{kern}:{slots}"
                results[key_ks] = max(results.get(key_ks, 0.0), round(scale / el / 1e9, 2))
                if log:
                    log(f"autotune kernel {kern} slots {slots}: {round(scale / el / 1e9, 2)} Gcell/s")
    if not results:
        raise MemoryError("no slot count fits the device memory budget")
    best = max(results, key=results.get)
    kern, slots = (int(v) for v in best.split(":"))
    entry = {"kernel": kern, "slots": slots, "measured_gcell_s": results,
             "nt_scale": nt_scale}
    if write_cache:
        cache[key] = entry
        with open(CACHE, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    return entry
