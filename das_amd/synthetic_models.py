"""Synthetic fault benchmark models (tsw_rtm/synthetic_models.py:16-164).

Same names, parameters and returned fields as the reference; the models are
built row-vectorised on the host (they are inputs, built once, not part of
the propagation).  Fault columns follow the reference's rounding
(round-half-even of surface index + z cot(dip) / dx, clipped to the grid).
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class SyntheticBenchmark:
    name: str
    velocity_model: np.ndarray  # (nz, nx)
    fault_mask: np.ndarray      # binary mask of true fault plane
    nx: int
    nz: int
    dx: float
    dz: float
    fault_x_surface: float
    v_sediment_surface: float
    v_rock: float


def _fault_columns(x0, z, cot, dx, nx):
    return np.clip(np.rint(x0 + (z * cot) / dx).astype(np.int64), 0, nx - 1)


def _build(nx, nz, dz, v_sed_top, v_sed_grad, v_rock, rock_lo, rock_hi, faults):
    """Rock where rock_lo[iz] <= ix < rock_hi[iz], sediment (v_sed_top +
    v_sed_grad * z) elsewhere; fault_mask set at the given columns per row."""
    z = np.arange(nz) * dz
    cols = np.arange(nx)[None, :]
    v_sed = (v_sed_top + v_sed_grad * z)[:, None]
    rock = (cols >= rock_lo[:, None]) & (cols < rock_hi[:, None])
    vel = np.where(rock, v_rock, v_sed).astype(np.float32)
    mask = np.zeros((nz, nx), bool)
    for fx in faults:
        mask[np.arange(nz), fx] = True
    return vel, mask


def create_vertical_fault_model(nx: int = 300, nz: int = 150, dx: float = 20.0, dz: float = 20.0,
                                fault_x_idx: int = 150, v_sed_top: float = 1200.0,
                                v_sed_grad: float = 0.2, v_rock: float = 3000.0
                                ) -> SyntheticBenchmark:
    """synthetic_models.py:30-66: vertical fault, sediment (with a velocity
    gradient) left of `fault_x_idx`, rock from it on."""
    lo = np.full(nz, fault_x_idx)
    vel, mask = _build(nx, nz, dz, v_sed_top, v_sed_grad, v_rock, lo, np.full(nz, 1 << 62),
                       [lo])
    return SyntheticBenchmark("Vertical Fault", vel, mask, nx, nz, dx, dz, fault_x_idx * dx,
                              v_sed_top, v_rock)


def create_dipping_fault_model(dip_angle_deg: float = 65.0, nx: int = 300, nz: int = 150,
                               dx: float = 20.0, dz: float = 20.0, surface_fault_x_idx: int = 150,
                               v_sed_top: float = 1200.0, v_sed_grad: float = 0.2,
                               v_rock: float = 3000.0) -> SyntheticBenchmark:
    """synthetic_models.py:69-115: fault trace x(z) = x0 + z cot(dip)."""
    z = np.arange(nz) * dz
    fx = _fault_columns(surface_fault_x_idx, z, 1.0 / np.tan(np.radians(dip_angle_deg)), dx, nx)
    vel, mask = _build(nx, nz, dz, v_sed_top, v_sed_grad, v_rock, fx, np.full(nz, 1 << 62), [fx])
    return SyntheticBenchmark(f"Dipping Fault ({dip_angle_deg:.0f} deg)", vel, mask, nx, nz, dx,
                              dz, surface_fault_x_idx * dx, v_sed_top, v_rock)


def create_fault_zone_model(nx: int = 350, nz: int = 150, dx: float = 20.0, dz: float = 20.0,
                            fault_left_x_idx: int = 110, fault_right_x_idx: int = 240,
                            v_sed_top: float = 1200.0, v_sed_grad: float = 0.2,
                            v_rock: float = 3000.0) -> SyntheticBenchmark:
    """synthetic_models.py:118-164: a rock block between two faults dipping
    outwards at 70 degrees, sediment on both sides."""
    z = np.arange(nz) * dz
    c = 1.0 / np.tan(np.radians(70.0))
    fl = _fault_columns(fault_left_x_idx, z, -c, dx, nx)
    fr = _fault_columns(fault_right_x_idx, z, c, dx, nx)
    vel, mask = _build(nx, nz, dz, v_sed_top, v_sed_grad, v_rock, fl, fr, [fl, fr])
    return SyntheticBenchmark("Dual Fault Zone (TFZ Style)", vel, mask, nx, nz, dx, dz,
                              fault_left_x_idx * dx, v_sed_top, v_rock)
