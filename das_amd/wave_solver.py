"""Drop-in for tsw_rtm.wave_solver on MI355X.

Same public names, signatures and return types as tsw_rtm/wave_solver.py
(``ricker_wavelet``, ``WaveSolver2D`` and its methods), but every time loop
runs on the GPU through libtswrtm_hip.so: the per-step source injection,
receiver recording, snapshot capture and imaging that the reference performs
in Python around each C call are fused into the HIP step kernel, and one
ctypes call covers a whole propagation.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _native as N


def ricker_wavelet(f0: float, dt: float, t_length: float) -> Tuple[np.ndarray, np.ndarray]:
    """Zero-phase Ricker wavelet (tsw_rtm/wave_solver.py:74-81): FP64 closed
    form, ``nt = round(t_length/dt)``, peak at t = 1/f0, returned as float32."""
    nt = int(round(t_length / dt))
    t = np.arange(nt) * dt
    arg = (np.pi * f0 * (t - 1.0 / f0)) ** 2
    w = (1.0 - 2.0 * arg) * np.exp(-arg)
    return w.astype(np.float32), t


def _norm_index(i, n, what):
    """numpy indexing semantics of the reference: negatives wrap, others raise."""
    i = int(i)
    j = i + n if i < 0 else i
    if not 0 <= j < n:
        raise IndexError(f"{what} {i} is out of bounds for axis with size {n}")
    return j


def _norm_indices(a, n, what):
    a = np.asarray(a).astype(np.int64).reshape(-1)
    b = np.where(a < 0, a + n, a)
    bad = (b < 0) | (b >= n)
    if bad.any():
        raise IndexError(f"{what} {a[bad][0]} is out of bounds for axis with size {n}")
    return N.i32(b)


class WaveSolver2D:
    """2D wavefield propagator on the HIP engine (tsw_rtm/wave_solver.py:84-242).

    ``sponge`` is the host copy of the Cerjan sponge (built on the device,
    bit-identical to wave2d.c:237-265).  As in the reference, the array is read
    again by every propagation: an in-place modification is uploaded before
    the next call.
    """

    def __init__(self, nx: int, nz: int, dx: float, dz: float, dt: float, n_sponge: int = 30,
                 sponge_alpha: float = 0.0015, free_surface_top: bool = True,
                 solver_type: str = "acoustic"):
        self.nx = nx
        self.nz = nz
        self.dx = float(dx)
        self.dz = float(dz)
        self.dt = float(dt)
        self.n_sponge = n_sponge
        self.sponge_alpha = float(sponge_alpha)
        self.free_surface_top = free_surface_top
        self.solver_type = solver_type
        self._grid = None
        h = N.P()
        # the reference passes alpha through c_float (wave_solver.py:113-116)
        alpha32 = float(np.float32(sponge_alpha))
        N.check(N.lib().tsw_grid_create(int(nx), int(nz), self.dx, self.dz, self.dt,
                                        int(n_sponge), alpha32,
                                        1 if free_surface_top else 0, ctypes.byref(h)))
        self._grid = h
        self.sponge = np.ones((nz, nx), dtype=np.float32)
        N.check(N.lib().tsw_grid_sponge(self._grid, N.ptr(self.sponge)))
        self._sponge_on_device = self.sponge.copy()

    def __del__(self):
        g = getattr(self, "_grid", None)
        if g is not None and N._lib is not None:
            try:
                N.lib().tsw_grid_destroy(g)
            except Exception:
                pass
            self._grid = None

    def _sync_sponge(self):
        sp = N.f32(self.sponge)
        if sp.shape != (self.nz, self.nx):
            raise ValueError(f"sponge has shape {sp.shape}, expected {(self.nz, self.nx)}")
        if not np.array_equal(sp, self._sponge_on_device, equal_nan=True):
            N.check(N.lib().tsw_grid_set_sponge(self._grid, N.ptr(sp)))
            self._sponge_on_device = sp.copy()

    def _model(self, velocity_model):
        vel = N.f32(velocity_model)
        if vel.shape != (self.nz, self.nx):
            raise ValueError(f"velocity model has shape {vel.shape}, expected {(self.nz, self.nx)}")
        return vel

    def check_stability(self, max_vel: float, min_vel: float, max_freq: float) -> Tuple[bool, str]:
        """CFL / points-per-wavelength check, as tsw_rtm/wave_solver.py:119-126
        (message thresholds 0.6 / 5, tested thresholds 0.65 / 4.0)."""
        cfl = max_vel * self.dt * np.sqrt(1.0 / (self.dx ** 2) + 1.0 / (self.dz ** 2))
        grid_pts_per_wl = (min_vel / max_freq) / max(self.dx, self.dz)
        msg = (f"CFL number: {cfl:.3f} (target < 0.6). "
               f"Grid pts per min wavelength: {grid_pts_per_wl:.1f} (target >= 5).")
        stable = (cfl < 0.65) and (grid_pts_per_wl >= 4.0)
        return stable, msg

    def forward_propagate_acoustic(self, velocity_model: np.ndarray, source_ts: np.ndarray,
                                   source_ix: int, source_iz: int, receiver_ix: np.ndarray,
                                   receiver_iz: int = 2, save_wavefield: bool = False,
                                   snapshot_interval: int = 1
                                   ) -> Tuple[np.ndarray, Optional[np.ndarray]]:
        """Forward modelling (tsw_rtm/wave_solver.py:128-186).

        Returns ``(recorded_traces (nt, n_rec), wavefields (n_snaps, nz, nx) or None)``.
        Snapshot j is p_curr after injection and before the step at it = j*k.
        """
        vel = self._model(velocity_model)
        src = N.f32(np.asarray(source_ts).reshape(-1))
        nt = len(src)
        six = _norm_index(source_ix, self.nx, "source_ix")
        siz = _norm_index(source_iz, self.nz, "source_iz")
        rix = _norm_indices(receiver_ix, self.nx, "receiver_ix")
        riz = _norm_index(receiver_iz, self.nz, "receiver_iz")
        k = int(snapshot_interval)
        if k < 1:
            raise ValueError("snapshot_interval must be >= 1")
        n_rec = len(rix)
        rec = np.zeros((nt, n_rec), dtype=np.float32)
        if nt == 0:
            return rec, (np.zeros((0, self.nz, self.nx), np.float32) if save_wavefield else None)
        snaps = (np.empty(((nt - 1) // k + 1, self.nz, self.nx), dtype=np.float32)
                 if save_wavefield else None)
        self._sync_sponge()
        N.check(N.lib().tsw_forward_acoustic(self._grid, N.ptr(vel), N.ptr(src), nt, six, siz,
                                             N.ptr(rix), n_rec, riz, k, N.ptr(rec), N.ptr(snaps)))
        return rec, snaps

    def backward_propagate_and_correlate(self, velocity_model: np.ndarray,
                                         receiver_data: np.ndarray, receiver_ix: np.ndarray,
                                         receiver_iz: int, forward_wavefields: np.ndarray,
                                         image_out: np.ndarray, snapshot_interval: int = 1):
        """Time-reversed receiver injection + zero-lag imaging
        (tsw_rtm/wave_solver.py:188-242); accumulates into ``image_out`` in place.

        The snapshot cursor starts at (nt-1)//k and only moves while it lies
        inside ``forward_wavefields``, as in the reference.
        """
        vel = self._model(velocity_model)
        data = N.f32(receiver_data)
        nt, n_rec = data.shape
        rix = _norm_indices(receiver_ix, self.nx, "receiver_ix")
        if len(rix) != n_rec:
            raise ValueError(f"{len(rix)} receiver positions for {n_rec} traces")
        riz = _norm_index(receiver_iz, self.nz, "receiver_iz")
        k = int(snapshot_interval)
        if k < 1:
            raise ValueError("snapshot_interval must be >= 1")
        if nt == 0:
            return
        snaps = N.f32(forward_wavefields)
        if snaps.ndim != 3 or snaps.shape[1:] != (self.nz, self.nx):
            raise ValueError(f"forward_wavefields has shape {snaps.shape}")
        if not (isinstance(image_out, np.ndarray) and image_out.shape == (self.nz, self.nx)):
            raise ValueError("image_out must be an (nz, nx) array")
        img = N.f32(image_out)
        self._sync_sponge()
        N.check(N.lib().tsw_backward_correlate(self._grid, N.ptr(vel), N.ptr(data), nt,
                                               N.ptr(rix), n_rec, riz, N.ptr(snaps),
                                               len(snaps), N.ptr(img), k))
        if img is not image_out:
            image_out[...] = img

    # ------------------------------------------------------------- elastic
    # The reference declares elastic_2d_step (wave_solver.py:36-53) but never
    # drives it; solver_type="elastic" is its only hook.  These methods are the
    # elastic counterparts of the acoustic loops, with the conventions of
    # include/tswrtm.h (vertical force source, vertical strain-rate receivers,
    # vz imaging).

    def _elastic_model(self, model):
        if len(model) != 4:
            raise ValueError("elastic model is (lambda, mu, buoyancy_x, buoyancy_z)")
        arrs = [N.f32(a) for a in model]
        for a in arrs:
            if a.shape != (self.nz, self.nx):
                raise ValueError(f"elastic model array has shape {a.shape}")
        return arrs

    def forward_propagate_elastic(self, model, source_ts, source_ix, source_iz, receiver_ix,
                                  receiver_iz=2, save_wavefield=False, snapshot_interval=1,
                                  return_illumination=False):
        """Elastic forward modelling -> (strain-rate traces (nt, n_rec),
        vz snapshots or None[, sum of squared snapshots])."""
        par = self._elastic_model(model)
        src = N.f32(np.asarray(source_ts).reshape(-1))
        nt = len(src)
        six = _norm_index(source_ix, self.nx, "source_ix")
        siz = _norm_index(source_iz, self.nz, "source_iz")
        rix = _norm_indices(receiver_ix, self.nx, "receiver_ix")
        riz = _norm_index(receiver_iz, self.nz, "receiver_iz")
        k = int(snapshot_interval)
        if k < 1 or nt == 0:
            raise ValueError("need snapshot_interval >= 1 and a non-empty source")
        rec = np.zeros((nt, len(rix)), np.float32)
        snaps = (np.empty(((nt - 1) // k + 1, self.nz, self.nx), np.float32)
                 if save_wavefield else None)
        ill = np.empty((self.nz, self.nx), np.float32) if return_illumination else None
        self._sync_sponge()
        N.check(N.lib().tsw_forward_elastic(self._grid, *[N.ptr(a) for a in par], N.ptr(src), nt,
                                            six, siz, N.ptr(rix), len(rix), riz, k, N.ptr(rec),
                                            N.ptr(snaps), N.ptr(ill)))
        return (rec, snaps, ill) if return_illumination else (rec, snaps)

    def backward_propagate_and_correlate_elastic(self, model, receiver_data, receiver_ix,
                                                 receiver_iz, forward_wavefields, image_out,
                                                 snapshot_interval=1):
        """Elastic back-propagation of the data + vz imaging into image_out."""
        par = self._elastic_model(model)
        data = N.f32(receiver_data)
        nt, n_rec = data.shape
        rix = _norm_indices(receiver_ix, self.nx, "receiver_ix")
        if len(rix) != n_rec:
            raise ValueError(f"{len(rix)} receiver positions for {n_rec} traces")
        riz = _norm_index(receiver_iz, self.nz, "receiver_iz")
        k = int(snapshot_interval)
        snaps = N.f32(forward_wavefields)
        img = N.f32(image_out)
        self._sync_sponge()
        N.check(N.lib().tsw_backward_correlate_elastic(
            self._grid, *[N.ptr(a) for a in par], N.ptr(data), nt, N.ptr(rix), n_rec, riz,
            N.ptr(snaps), len(snaps), N.ptr(img), k))
        if img is not image_out:
            image_out[...] = img


def elastic_parameters(vp, vs, rho):
    """(lambda, mu, buoyancy_x, buoyancy_z) float32 from Vp, Vs and density,
    with cell-centred buoyancy 1/rho on both staggered velocity grids."""
    vp = np.asarray(vp, np.float64)
    vs = np.asarray(vs, np.float64)
    rho = np.asarray(rho, np.float64)
    mu = rho * vs * vs
    lam = rho * vp * vp - 2.0 * mu
    b = 1.0 / rho
    return tuple(np.ascontiguousarray(a, np.float32) for a in (lam, mu, b, b))
