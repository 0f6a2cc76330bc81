"""Velocity flood models for the RTM engine.

``FloodModels`` has the fields of tsw_rtm/velocity_flood.py:15-21; the RTM
path reads ``sediment_flood`` (forward model) and ``rock_flood``
(back-propagation model), rtm_engine.py:78 and :95.
``generate_velocity_flood_models`` (velocity_flood.py:24-94) builds them on
the device: the column each flood column copies is worked out on the host by
the reference's own slice and loop assignments applied to column indices
(so offsets past the grid edge behave as there), and a HIP kernel gathers
and scales the columns.
"""
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class FloodModels:
    true_model: np.ndarray
    sediment_flood: np.ndarray
    rock_flood: np.ndarray
    flood_line_sed: int
    flood_line_rock: int


def flood_models_from(migration_model: np.ndarray, rock_model: np.ndarray = None) -> FloodModels:
    """FloodModels with one migration model used for both propagations
    (the single-model RTM case), or distinct forward / backward models."""
    m = np.ascontiguousarray(migration_model, dtype=np.float32)
    r = m if rock_model is None else np.ascontiguousarray(rock_model, dtype=np.float32)
    return FloodModels(true_model=m, sediment_flood=m, rock_flood=r, flood_line_sed=0,
                       flood_line_rock=0)


def _column_map(nx, start, keep_left):
    """Column indices left by `m[:, :start] = b[:, :start]` then
    `for ix in range(start, nx): m[:, ix] = b[:, start]` (keep_left), or by
    `m[:, start:] = b[:, start:]` then `for ix in range(0, start): m[:, ix] =
    b[:, start]`, evaluated with numpy's indexing rules on an index vector."""
    cols = np.arange(nx)
    cm = np.full(nx, -1, np.int64)
    if keep_left:
        cm[:start] = cols[:start]
        for ix in range(start, nx):
            cm[ix] = cols[start]
    else:
        cm[start:] = cols[start:]
        for ix in range(0, start):
            cm[ix] = cols[start]
    return np.ascontiguousarray(cm, dtype=np.int32)


def _flood(base, cm, perturb):
    nz, nx = base.shape
    out = np.empty((nz, nx), np.float32)
    N.check(N.lib().tsw_flood_columns(N.ptr(base), base.dtype.itemsize, nx, nz, N.ptr(cm),
                                      1 if perturb != 0.0 else 0, 1.0 + perturb, N.ptr(out)))
    return out


def generate_velocity_flood_models(base_model: np.ndarray, fault_x_idx: int,
                                   sediment_is_left: bool = True, flood_offset_cells: int = 15,
                                   perturb_sediment: float = 0.0,
                                   perturb_rock: float = 0.0) -> FloodModels:
    """velocity_flood.py:24-94: the sediment flood extends the sediment-side
    profile at `flood_offset_cells` from the fault across the rock side, the
    rock flood the rock-side profile across the sediment side; optional
    fractional perturbations (e.g. +/-10%)."""
    b = np.asarray(base_model)
    if b.ndim != 2:
        raise ValueError("base_model must be (nz, nx)")
    nz, nx = b.shape
    work = np.ascontiguousarray(b, dtype=np.float64 if b.dtype == np.float64 else np.float32)
    if sediment_is_left:
        fx_sed = min(nx - 1, fault_x_idx - flood_offset_cells)
        fx_rock = max(0, fault_x_idx + flood_offset_cells)
        cm_sed = _column_map(nx, fx_sed, True)
        cm_rock = _column_map(nx, fx_rock, False)
    else:
        fx_sed = max(0, fault_x_idx + flood_offset_cells)
        fx_rock = min(nx - 1, fault_x_idx - flood_offset_cells)
        cm_sed = _column_map(nx, fx_sed, False)
        cm_rock = _column_map(nx, fx_rock, True)
    return FloodModels(true_model=b.copy(), sediment_flood=_flood(work, cm_sed, perturb_sediment),
                       rock_flood=_flood(work, cm_rock, perturb_rock), flood_line_sed=fx_sed,
                       flood_line_rock=fx_rock)
