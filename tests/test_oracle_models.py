"""Model-side builders: the oracle's flood models are pinned by their
defining properties (profiles copied up to the flood line, the flood-line
column repeated beyond it, perturbation scaling, the flood lines themselves,
offsets past the grid edges); the synthetic fault models by the fault column
each row implies from the dip angle; the host builders of das_amd equal the
oracle bit for bit (the flood models' device path is checked in
tests/test_gpu_models.py)."""
import numpy as np
import pytest

from oracle import model_oracle as MO


def _base(nz=20, nx=60, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    return (1500 + 1000 * rng.random((nz, nx))).astype(dtype)


@pytest.mark.parametrize("left", [True, False])
def test_flood_profiles(left):
    b = _base()
    sed, rock, fs, fr = MO.flood_models(b, 30, left, 10)
    if left:
        assert (fs, fr) == (20, 40)
        np.testing.assert_array_equal(sed[:, :20], b[:, :20])
        np.testing.assert_array_equal(sed[:, 20:], np.repeat(b[:, 20:21], 40, axis=1))
        np.testing.assert_array_equal(rock[:, 40:], b[:, 40:])
        np.testing.assert_array_equal(rock[:, :40], np.repeat(b[:, 40:41], 40, axis=1))
    else:
        assert (fs, fr) == (40, 20)
        np.testing.assert_array_equal(sed[:, 40:], b[:, 40:])
        np.testing.assert_array_equal(sed[:, :40], np.repeat(b[:, 40:41], 40, axis=1))
        np.testing.assert_array_equal(rock[:, :20], b[:, :20])
        np.testing.assert_array_equal(rock[:, 20:], np.repeat(b[:, 20:21], 40, axis=1))


def test_flood_perturbation_and_edges():
    b = _base()
    sed, rock, _, _ = MO.flood_models(b, 30, True, 10, 0.1, -0.1)
    s0, r0, _, _ = MO.flood_models(b, 30, True, 10)
    np.testing.assert_array_equal(sed, (s0 * np.float32(1.1)).astype(np.float32))
    np.testing.assert_array_equal(rock, (r0 * np.float32(0.9)).astype(np.float32))
    # flood line before the first column: numpy's negative index picks a column
    # from the right, and the loop then fills every column with it
    sed, _, fs, _ = MO.flood_models(b, 3, True, 10)
    assert fs == -7
    np.testing.assert_array_equal(sed, np.repeat(b[:, 53:54], 60, axis=1))


def test_synthetic_fault_columns():
    vel, mask = MO.dipping_fault(45.0, nx=100, nz=30, dx=10.0, dz=10.0, surface_fault_x_idx=20)
    # 45 degrees, dx = dz: the fault moves one column per row
    for iz in range(30):
        fx = int(np.flatnonzero(mask[iz])[0])
        assert fx == 20 + iz
        assert (vel[iz, fx:] == 3000.0).all() and (vel[iz, :fx] < 3000.0).all()
    vel, mask = MO.vertical_fault(nx=40, nz=10, fault_x_idx=13)
    assert mask[:, 13].all() and mask.sum() == 10
    np.testing.assert_array_equal(vel[:, 0], np.float32(1200.0 + 0.2 * np.arange(10) * 20.0))
    vel, mask = MO.fault_zone(nx=100, nz=40, dx=10.0, dz=10.0, fault_left_x_idx=30,
                              fault_right_x_idx=70)
    assert mask.sum(axis=1).max() == 2
    assert (vel[0, 30:70] == 3000.0).all() and (vel[0, :30] < 3000.0).all()


@pytest.mark.parametrize("kw", [{}, {"dip_angle_deg": 35.0, "surface_fault_x_idx": 130},
                                {"dip_angle_deg": 89.0, "nx": 50, "surface_fault_x_idx": 49}])
def test_host_builders_equal_oracle(kw):
    from das_amd import synthetic_models as SM

    m = SM.create_dipping_fault_model(**kw)
    vel, mask = MO.dipping_fault(**kw)
    np.testing.assert_array_equal(m.velocity_model, vel)
    np.testing.assert_array_equal(m.fault_mask, mask)
    m = SM.create_vertical_fault_model()
    vel, mask = MO.vertical_fault()
    np.testing.assert_array_equal(m.velocity_model, vel)
    np.testing.assert_array_equal(m.fault_mask, mask)
    for zkw in ({}, {"fault_left_x_idx": 200, "fault_right_x_idx": 100}):
        m = SM.create_fault_zone_model(**zkw)
        vel, mask = MO.fault_zone(**zkw)
        np.testing.assert_array_equal(m.velocity_model, vel)
        np.testing.assert_array_equal(m.fault_mask, mask)
