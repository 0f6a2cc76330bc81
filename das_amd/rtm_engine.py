"""Drop-in for tsw_rtm.rtm_engine on MI355X.

Same public names and return types as tsw_rtm/rtm_engine.py (``RTMResult``,
``apply_laplacian_filter``, ``TSWRTMEngine``).  ``migrate_shot`` keeps the
forward snapshots in HBM (the reference holds them in a Python list and sums
their squares on the host); ``migrate_multiple_shots`` also stacks, mutes,
normalises and filters on the device.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from . import _native as N
from .velocity_flood import FloodModels
from .wave_solver import WaveSolver2D, _norm_index, _norm_indices, ricker_wavelet  # noqa: F401


@dataclass
class RTMResult:
    raw_image: np.ndarray
    laplacian_image: np.ndarray
    normalized_image: np.ndarray
    grid_x: np.ndarray
    grid_z: np.ndarray
    dx: float
    dz: float


def apply_laplacian_filter(image: np.ndarray, dx: float, dz: float) -> np.ndarray:
    """5-point Laplacian with symmetric boundary (tsw_rtm/rtm_engine.py:27-38),
    computed on the GPU, bit-identical to the reference's scipy call."""
    img = N.f32(image)
    if img.ndim != 2:
        raise ValueError("image must be 2-D")
    out = np.empty_like(img)
    if img.size:
        N.check(N.lib().tsw_laplacian(N.ptr(img), N.ptr(out), img.shape[1], img.shape[0],
                                      float(dx), float(dz)))
    return out


def _check_kept(status, *arrays):
    """N.check, except that a diverged shot's finished arrays (NaN/Inf in
    them) ride on the exception as ``result`` for the caller to inspect."""
    try:
        N.check(status)
    except N.WavefieldDiverged as e:
        e.result = arrays
        raise


class TSWRTMEngine:
    """Transmitted-surface-wave RTM engine (tsw_rtm/rtm_engine.py:41-171)."""

    def __init__(self, solver: WaveSolver2D):
        self.solver = solver

    def _shot_args(self, source_ts, source_ix, source_iz, receiver_data, receiver_ix,
                   receiver_iz, snapshot_interval):
        s = self.solver
        src = N.f32(np.asarray(source_ts).reshape(-1))
        data = N.f32(receiver_data)
        if data.ndim != 2:
            raise ValueError("receiver_data must be (nt, n_receivers)")
        rix = _norm_indices(receiver_ix, s.nx, "receiver_ix")
        if len(rix) != data.shape[1]:
            raise ValueError(f"{len(rix)} receiver positions for {data.shape[1]} traces")
        six = _norm_index(source_ix, s.nx, "source_ix")
        siz = _norm_index(source_iz, s.nz, "source_iz")
        riz = _norm_index(receiver_iz, s.nz, "receiver_iz")
        k = int(snapshot_interval)
        if k < 1:
            raise ValueError("snapshot_interval must be >= 1")
        if len(src) == 0 or data.shape[0] == 0:
            raise ValueError("empty source or receiver data")
        if len(src) != data.shape[0]:
            # the reference would run the forward over len(source_ts) and the
            # back-propagation over receiver_data.shape[0], pairing shifted
            # snapshots without notice (rtm_engine.py:76-104)
            raise ValueError(f"source_ts has {len(src)} samples but receiver_data has "
                             f"{data.shape[0]}; they must be equal")
        return src, six, siz, data, rix, riz, k

    def migrate_shot(self, flood_models: FloodModels, source_ts: np.ndarray, source_ix: int,
                     source_iz: int, receiver_data: np.ndarray, receiver_ix: np.ndarray,
                     receiver_iz: int = 2, snapshot_interval: int = 2
                     ) -> Tuple[np.ndarray, np.ndarray]:
        """One virtual shot (tsw_rtm/rtm_engine.py:49-104) -> (shot_image, illumination).

        Forward in ``flood_models.sediment_flood`` with snapshots every k steps
        kept in HBM; illumination = sum of squared snapshots * dt*k; then
        back-propagation of ``receiver_data`` in ``flood_models.rock_flood``.
        """
        s = self.solver
        src, six, siz, data, rix, riz, k = self._shot_args(
            source_ts, source_ix, source_iz, receiver_data, receiver_ix, receiver_iz,
            snapshot_interval)
        vf = s._model(flood_models.sediment_flood)
        vb = s._model(flood_models.rock_flood)
        img = np.empty((s.nz, s.nx), np.float32)
        ill = np.empty((s.nz, s.nx), np.float32)
        s._sync_sponge()
        _check_kept(N.lib().tsw_migrate_shot(s._grid, N.ptr(vf), N.ptr(vb), N.ptr(src),
                                             len(src), six, siz, N.ptr(data), data.shape[0],
                                             N.ptr(rix), len(rix), riz, k, N.ptr(img),
                                             N.ptr(ill)), img, ill)
        return img, ill

    def migrate_shot_elastic(self, model_fwd, model_bwd, source_ts, source_ix, source_iz,
                             receiver_data, receiver_ix, receiver_iz=2, snapshot_interval=2):
        """Elastic counterpart of migrate_shot: forward in ``model_fwd``
        (lambda, mu, buoyancy_x, buoyancy_z) with vz snapshots in HBM, back-
        propagation of the strain-rate gather in ``model_bwd`` -> (image, illumination)."""
        s = self.solver
        src, six, siz, data, rix, riz, k = self._shot_args(
            source_ts, source_ix, source_iz, receiver_data, receiver_ix, receiver_iz,
            snapshot_interval)
        mf = s._elastic_model(model_fwd)
        mb = s._elastic_model(model_bwd)
        PF = (N.P * 4)(*[N.ptr(a) for a in mf])
        PB = (N.P * 4)(*[N.ptr(a) for a in mb])
        img = np.empty((s.nz, s.nx), np.float32)
        ill = np.empty_like(img)
        s._sync_sponge()
        _check_kept(N.lib().tsw_migrate_shot_elastic(s._grid, PF, PB, N.ptr(src), six, siz,
                                                     N.ptr(data), data.shape[0], N.ptr(rix),
                                                     len(rix), riz, k, N.ptr(img), N.ptr(ill)),
                    img, ill)
        return img, ill

    def migrate_multiple_shots(self, flood_models: FloodModels, source_waveforms: List[np.ndarray],
                               source_coords_ix: List[int], source_coords_iz: List[int],
                               receiver_gathers: List[np.ndarray], receiver_ix: np.ndarray,
                               receiver_iz: int = 2, snapshot_interval: int = 2,
                               mute_sponge: bool = True) -> RTMResult:
        """Multi-shot RTM with stacking and post-processing
        (tsw_rtm/rtm_engine.py:106-171), all on the GPU."""
        s = self.solver
        n_shots = len(source_coords_ix)
        shots = [self._shot_args(source_waveforms[i], source_coords_ix[i], source_coords_iz[i],
                                 receiver_gathers[i], receiver_ix, receiver_iz, snapshot_interval)
                 for i in range(n_shots)]
        rix = shots[0][4] if shots else _norm_indices(receiver_ix, s.nx, "receiver_ix")
        riz = shots[0][5] if shots else _norm_index(receiver_iz, s.nz, "receiver_iz")
        k = int(snapshot_interval)
        vf = s._model(flood_models.sediment_flood)
        vb = s._model(flood_models.rock_flood)
        PA = N.P * max(n_shots, 1)
        IA = N.I * max(n_shots, 1)
        srcs = PA(*[N.ptr(sh[0]) for sh in shots])
        datas = PA(*[N.ptr(sh[3]) for sh in shots])
        nt_src = IA(*[len(sh[0]) for sh in shots])
        nts = IA(*[sh[3].shape[0] for sh in shots])
        sixs = IA(*[sh[1] for sh in shots])
        sizs = IA(*[sh[2] for sh in shots])
        raw = np.empty((s.nz, s.nx), np.float32)
        lap = np.empty_like(raw)
        norm = np.empty_like(raw)
        s._sync_sponge()
        _check_kept(N.lib().tsw_migrate_shots(s._grid, N.ptr(vf), N.ptr(vb), n_shots, srcs,
                                              nt_src, sixs, sizs, datas, nts, N.ptr(rix),
                                              len(rix), riz, k, 1 if mute_sponge else 0,
                                              N.ptr(raw), N.ptr(lap), N.ptr(norm)),
                    raw, lap, norm)
        return RTMResult(raw_image=raw, laplacian_image=lap, normalized_image=norm,
                         grid_x=np.arange(s.nx) * s.dx, grid_z=np.arange(s.nz) * s.dz,
                         dx=s.dx, dz=s.dz)
