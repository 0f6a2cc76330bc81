"""CPU oracle for the model-side builders.

TEST INFRASTRUCTURE ONLY: the checker das_amd.velocity_flood and
das_amd.synthetic_models are compared against; the product never imports it.

Restates, column by column and row by row as the reference loops do:
  - generate_velocity_flood_models   tsw_rtm/velocity_flood.py:24-94
  - create_vertical_fault_model      tsw_rtm/synthetic_models.py:30-66
  - create_dipping_fault_model       tsw_rtm/synthetic_models.py:69-115
  - create_fault_zone_model          tsw_rtm/synthetic_models.py:118-164
Pinned in tests/test_oracle_models.py by invariants (profiles, flood lines,
fault positions from the dip angle).
"""
import numpy as np


def flood_models(base, fault_x_idx, sediment_is_left=True, offset=15, p_sed=0.0, p_rock=0.0):
    """-> (sediment_flood, rock_flood, flood_line_sed, flood_line_rock)."""
    nz, nx = base.shape
    sed = np.zeros_like(base)
    rock = np.zeros_like(base)
    if sediment_is_left:
        fs = min(nx - 1, fault_x_idx - offset)
        fr = max(0, fault_x_idx + offset)
        sed[:, :fs] = base[:, :fs]
        for ix in range(fs, nx):
            sed[:, ix] = base[:, fs]
        rock[:, fr:] = base[:, fr:]
        for ix in range(0, fr):
            rock[:, ix] = base[:, fr]
    else:
        fs = max(0, fault_x_idx + offset)
        fr = min(nx - 1, fault_x_idx - offset)
        sed[:, fs:] = base[:, fs:]
        for ix in range(0, fs):
            sed[:, ix] = base[:, fs]
        rock[:, :fr] = base[:, :fr]
        for ix in range(fr, nx):
            rock[:, ix] = base[:, fr]
    if p_sed != 0.0:
        sed = sed * (1.0 + p_sed)
    if p_rock != 0.0:
        rock = rock * (1.0 + p_rock)
    return sed.astype(np.float32), rock.astype(np.float32), fs, fr


def vertical_fault(nx=300, nz=150, dx=20.0, dz=20.0, fault_x_idx=150, v_sed_top=1200.0,
                   v_sed_grad=0.2, v_rock=3000.0):
    vel = np.zeros((nz, nx), np.float32)
    mask = np.zeros((nz, nx), bool)
    for iz in range(nz):
        vel[iz, :fault_x_idx] = v_sed_top + v_sed_grad * (iz * dz)
        vel[iz, fault_x_idx:] = v_rock
        mask[iz, fault_x_idx] = True
    return vel, mask


def dipping_fault(dip_angle_deg=65.0, nx=300, nz=150, dx=20.0, dz=20.0, surface_fault_x_idx=150,
                  v_sed_top=1200.0, v_sed_grad=0.2, v_rock=3000.0):
    vel = np.zeros((nz, nx), np.float32)
    mask = np.zeros((nz, nx), bool)
    cot = 1.0 / np.tan(np.radians(dip_angle_deg))
    for iz in range(nz):
        z = iz * dz
        fx = max(0, min(nx - 1, int(round(surface_fault_x_idx + (z * cot) / dx))))
        vel[iz, :fx] = v_sed_top + v_sed_grad * z
        vel[iz, fx:] = v_rock
        mask[iz, fx] = True
    return vel, mask


def fault_zone(nx=350, nz=150, dx=20.0, dz=20.0, fault_left_x_idx=110, fault_right_x_idx=240,
               v_sed_top=1200.0, v_sed_grad=0.2, v_rock=3000.0):
    vel = np.zeros((nz, nx), np.float32)
    mask = np.zeros((nz, nx), bool)
    cl = -1.0 / np.tan(np.radians(70.0))
    cr = 1.0 / np.tan(np.radians(70.0))
    for iz in range(nz):
        z = iz * dz
        v_sed = v_sed_top + v_sed_grad * z
        fl = max(0, min(nx - 1, int(round(fault_left_x_idx + (z * cl) / dx))))
        fr = max(0, min(nx - 1, int(round(fault_right_x_idx + (z * cr) / dx))))
        vel[iz, :fl] = v_sed
        vel[iz, fl:fr] = v_rock
        vel[iz, fr:] = v_sed
        mask[iz, fl] = True
        mask[iz, fr] = True
    return vel, mask
