"""Flood models built on the device equal the oracle's restatement of
generate_velocity_flood_models bit for bit, for float32 and float64 base
models, both sediment sides, perturbations and flood lines past the grid."""
import numpy as np
import pytest

from oracle import model_oracle as MO

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("fault,off,ps,pr", [(60, 15, 0.0, 0.0), (60, 15, 0.1, -0.1),
                                             (5, 12, 0.0, 0.07), (108, 10, -0.03, 0.0)])
def test_flood_models_equal_oracle(dtype, left, fault, off, ps, pr):
    from das_amd.velocity_flood import generate_velocity_flood_models

    rng = np.random.default_rng(fault)
    base = (1500 + 1000 * rng.random((30, 120))).astype(dtype)
    fm = generate_velocity_flood_models(base, fault, left, off, ps, pr)
    sed, rock, fs, fr = MO.flood_models(base, fault, left, off, ps, pr)
    np.testing.assert_array_equal(fm.sediment_flood, sed)
    np.testing.assert_array_equal(fm.rock_flood, rock)
    assert (fm.flood_line_sed, fm.flood_line_rock) == (fs, fr)
    assert fm.sediment_flood.dtype == np.float32 and fm.true_model.dtype == dtype
    np.testing.assert_array_equal(fm.true_model, base)


def test_flood_line_past_the_right_edge_raises_like_numpy():
    from das_amd.velocity_flood import generate_velocity_flood_models

    base = np.ones((4, 20), np.float32)
    with pytest.raises(IndexError):
        MO.flood_models(base, 15, True, 10)
    with pytest.raises(IndexError):
        generate_velocity_flood_models(base, 15, True, 10)
