"""Batched multi-shot acoustic RTM on the HIP engine.

``ShotBatch`` uploads many virtual shots of uneven grid size and length once,
keeps their state in HBM and migrates them all with one kernel launch per
global time step (hipGraph-captured).  Per-shot images are bit-identical to
``TSWRTMEngine.migrate_shot`` (tsw_rtm/rtm_engine.py:49-104) on the same shot.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _native as N


def fill_desc(d, s, keep):
    """Fills tsw_shot_desc `d` from a shot object; arrays it points to are
    appended to `keep` (they must outlive the C call)."""
    src = N.f32(np.asarray(s.src).reshape(-1))
    data = N.f32(s.data)
    rix = N.i32(s.rix)
    if len(src) != s.nt or data.shape != (s.nt, len(rix)):
        raise ValueError("source / data length must equal the shot's nt")
    keep += [src, data, rix]
    if getattr(s, "kind", "acoustic") == "elastic":
        if s.model_fwd is None and s.model_bwd is None:
            d.kind = N.SHOT_ELASTIC  # resident models of a server: window at (z0, x0)
            mf = mb = None
        else:
            mf = [N.f32(a) for a in s.model_fwd]
            mb = [N.f32(a) for a in s.model_bwd]
        if mf is not None:
            if len(mf) != 4 or len(mb) != 4 or any(a.shape != (s.nz, s.nx) for a in mf + mb):
                raise ValueError("elastic model shape does not match the shot grid")
            keep += mf + mb
            d.kind = N.SHOT_ELASTIC
            for j in range(4):
                d.ela_fwd[j] = N.ptr(mf[j])
                d.ela_bwd[j] = N.ptr(mb[j])
    elif s.vel_fwd is None and s.vel_bwd is None:
        d.kind = N.SHOT_ACOUSTIC  # resident models of a server: window at (z0, x0)
    else:
        vf, vb = N.f32(s.vel_fwd), N.f32(s.vel_bwd)
        if vf.shape != (s.nz, s.nx) or vb.shape != (s.nz, s.nx):
            raise ValueError("model shape does not match the shot grid")
        keep += [vf, vb]
        d.kind = N.SHOT_ACOUSTIC
        d.vel_fwd, d.vel_bwd = N.ptr(vf), N.ptr(vb)
    d.nx, d.nz, d.nt = s.nx, s.nz, s.nt
    d.src = N.ptr(src)
    d.six, d.siz = int(s.six), int(s.siz)
    d.data, d.rix, d.nrec, d.riz = N.ptr(data), N.ptr(rix), len(rix), int(s.riz)
    d.x0, d.z0 = int(getattr(s, "x0", 0)), int(getattr(s, "z0", 0))


def kernel_configs():
    """{id: (T, W, H, threads)} of the temporally blocked kernel."""
    out = {}
    for i in range(1, N.lib().tsw_kernel_configs() + 1):
        v = [N.I() for _ in range(4)]
        N.check(N.lib().tsw_kernel_config(i, *[ctypes.byref(x) for x in v]))
        out[i] = tuple(x.value for x in v)
    return out


def kernel_name(kid: int) -> str:
    """Kernel and template arguments of blocked configuration `kid`, for
    example "k_tc<5,6>"."""
    buf = ctypes.create_string_buffer(64)
    N.check(N.lib().tsw_kernel_name(kid, buf, 64))
    return buf.value.decode()


class ShotBatch:
    """A resident batch of shots: objects with the fields of workloads.Shot
    (acoustic) or workloads.ElasticShot (kind == "elastic")."""

    def __init__(self, shots: Sequence, dx: float, dz: float, dt: float, n_sponge: int = 30,
                 sponge_alpha: float = 0.0015, free_surface_top: bool = True, k: int = 2,
                 slots: int = 0, graph_steps: int = 20000, kernel: int = 0):
        self.shots = list(shots)
        self.dx, self.dz, self.dt, self.k = float(dx), float(dz), float(dt), int(k)
        keep = []
        descs = (N.ShotDesc * len(self.shots))()
        for d, s in zip(descs, self.shots):
            fill_desc(d, s, keep)
        h = N.P()
        alpha32 = float(np.float32(sponge_alpha))
        N.check(N.lib().tsw_batch_create(len(self.shots), descs, self.dx, self.dz, self.dt,
                                         int(n_sponge), alpha32, 1 if free_surface_top else 0,
                                         self.k, int(slots), int(graph_steps), int(kernel),
                                         ctypes.byref(h)))
        self._h = h
        del keep  # uploaded synchronously by tsw_batch_create

    def close(self):
        if getattr(self, "_h", None) is not None:
            N.lib().tsw_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def run(self, mode: str = "batched"):
        m = {"batched": N.RUN_BATCHED, "plain": N.RUN_PLAIN}[mode]
        N.check(N.lib().tsw_batch_run(self._h, m))

    def time_kernels(self, lane: str = "acoustic"):
        """One run of a lane, timed by HIP events -> (milliseconds, launches).

        A temporally blocked batch with hipGraphs is timed as it runs: its
        graphs (each with the launch groups as parallel branches) inside one
        event pair; otherwise one event pair goes around each run of 32
        back-to-back launches.  A blocked launch is one block of T steps of
        every active shot."""
        ms, n = N.D(), N.I()
        N.check(N.lib().tsw_batch_time_kernels(self._h, 1 if lane == "elastic" else 0,
                                               ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def shot_result(self, i: int):
        s = self.shots[i]
        img = np.empty((s.nz, s.nx), np.float32)
        ill = np.empty_like(img)
        N.check(N.lib().tsw_batch_shot_result(self._h, int(i), N.ptr(img), N.ptr(ill)))
        return img, ill

    def stack(self, gnx: int, gnz: int):
        img = np.empty((gnz, gnx), np.float32)
        ill = np.empty_like(img)
        N.check(N.lib().tsw_batch_stack(self._h, int(gnx), int(gnz), N.ptr(img), N.ptr(ill)))
        return img, ill

    def stack_device(self, image, illum, stream=None):
        """Stacks into device tensors ``image`` and ``illum`` (float32, (gnz, gnx),
        contiguous) on ``stream`` (torch's current stream by default)."""
        import torch

        for t in (image, illum):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous() or t.dim() != 2:
                raise ValueError("stack_device needs contiguous 2-D float32 device tensors")
        if image.shape != illum.shape:
            raise ValueError("image and illumination shapes differ")
        gnz, gnx = image.shape
        st = (stream or torch.cuda.current_stream()).cuda_stream
        N.check(N.lib().tsw_batch_stack_device(self._h, int(gnx), int(gnz),
                                               ctypes.c_void_p(image.data_ptr()),
                                               ctypes.c_void_p(illum.data_ptr()),
                                               ctypes.c_void_p(st)))

    def info(self) -> dict:
        cu, nb, ecu = N.LL(), N.LL(), N.LL()
        steps, graphs = N.I(), N.I()
        N.check(N.lib().tsw_batch_info(self._h, ctypes.byref(cu), ctypes.byref(nb),
                                       ctypes.byref(steps), ctypes.byref(graphs),
                                       ctypes.byref(ecu)))
        return {"cell_updates": cu.value, "device_bytes": nb.value, "steps": steps.value,
                "graphs": graphs.value, "elastic_cell_updates": ecu.value}
