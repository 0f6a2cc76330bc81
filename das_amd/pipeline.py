"""The virtual-source RTM chain around the migration path, on the GPU.

Follows run_das_pipeline.py:160-370 step by step: per block of continuous
data, decimation, band-pass, 1-bit and whitening; correlation of each
virtual-source channel with the receiver channels, stacked over blocks and
divided by the block count; per source the causal Green's function (time
derivative) and an F-K fan filter; resampling to the simulation step; and
multi-shot migration with Ricker sources.  Each step is the das_amd function
of the same name as the reference's; per block, the preprocessing and the
correlation run back to back on the device (preprocess_block_device,
cross_correlations_device), so a block crosses PCIe once and the stack stays
in HBM until the last block.
"""
from __future__ import annotations

from typing import Iterable, List, Sequence, Tuple

import numpy as np

from .interferometry import build_virtual_shot_gather, cross_correlations_device, \
    resample_gather_to_sim_dt
from .preprocessing import fk_filter, preprocess_blocks_device
from .rtm_engine import RTMResult, TSWRTMEngine
from .velocity_flood import FloodModels
from .wave_solver import ricker_wavelet


def stacked_correlations(blocks: Iterable[np.ndarray], fs_orig: float, target_fs: float,
                         fmin: float, fmax: float, source_channels: Sequence[int],
                         receiver_channels: np.ndarray, max_lag_s: float,
                         group: int = 8) -> Tuple[np.ndarray, np.ndarray, float]:
    """run_das_pipeline.py:160-213 and :232-235 -> (ccfs (n_src, 2 n_lags + 1,
    n_rec) float32 averaged over the blocks, lags, dt_proc).  Up to ``group``
    consecutive blocks of one shape are preprocessed in one call (their
    channels filtered side by side); the correlations are added into the
    stack block by block, in order."""
    stacked, lags, dt_proc, n, whites = None, None, None, 0, None

    def flush(pending):
        nonlocal stacked, lags, dt_proc, n, whites
        # decimation, band-pass, 1-bit and whitening on the device; each block
        # crosses PCIe once and the stack stays resident until the end
        whites, dt_proc, _ = preprocess_blocks_device(pending, fs_orig, target_fs, fmin, fmax,
                                                      outs=whites)
        for w in whites[:len(pending)]:
            stacked, lags = cross_correlations_device(w, source_channels, max_lag_s, dt_proc,
                                                      receiver_channels, out=stacked)
            n += 1

    pending = []
    for raw in blocks:
        raw = np.asarray(raw)
        if pending and raw.shape != pending[0].shape:
            flush(pending)
            pending = []
        pending.append(raw)
        if len(pending) == max(1, int(group)):
            flush(pending)
            pending = []
    if pending:
        flush(pending)
    if stacked is None:
        raise ValueError("no data blocks")
    stacked = stacked.cpu().numpy()
    return stacked / np.float32(n) if n > 1 else stacked, lags, dt_proc


def virtual_gathers(ccfs: np.ndarray, lags: np.ndarray, source_channels: Sequence[int],
                    receiver_channels: np.ndarray, channel_coords_x: np.ndarray,
                    receiver_spacing_m: float, fk_vmin: float, fk_vmax: float) -> List[np.ndarray]:
    """run_das_pipeline.py:240-283: causal Green's function of each source's
    stacked correlations, then the F-K fan filter."""
    out = []
    for v_idx, src in enumerate(source_channels):
        vs = build_virtual_shot_gather(ccfs[v_idx], lags, v_idx, int(src), channel_coords_x,
                                       receiver_channels, branch="causal", take_derivative=True)
        out.append(fk_filter(vs.shot_gather, vs.dt, receiver_spacing_m, fk_vmin, fk_vmax))
    return out


def migrate_virtual_sources(engine: TSWRTMEngine, flood_models: FloodModels,
                            gathers: Sequence[np.ndarray], dt_proc: float, max_lag_s: float,
                            ricker_f0: float, source_ix: Sequence[int], receiver_ix: np.ndarray,
                            snapshot_interval: int = 2) -> RTMResult:
    """run_das_pipeline.py:337-370: gathers resampled to the solver's step,
    Ricker sources of max_lag_s, multi-shot migration from row 2."""
    dt_sim = engine.solver.dt
    sim_nt = int(round(max_lag_s / dt_sim))
    res = [resample_gather_to_sim_dt(g, dt_proc, dt_sim, sim_nt) for g in gathers]
    src, _ = ricker_wavelet(ricker_f0, dt_sim, max_lag_s)
    return engine.migrate_multiple_shots(flood_models, [src] * len(res), list(source_ix),
                                         [2] * len(res), res, receiver_ix, 2, snapshot_interval,
                                         mute_sponge=True)
