"""The native CSV writer (no GPU needed) produces the bytes of the
reference's save_matrix_csv formatting (restated in oracle/output_oracle.py):
float32 and float64 images, special values, integer and float axes.  The
oracle itself is pinned by parsing its output back."""
import numpy as np
import pytest

from oracle import output_oracle as OO


def _matrices():
    rng = np.random.default_rng(0)
    a = (rng.standard_normal((7, 11)) * 10.0 ** rng.integers(-40, 38, (7, 11))).astype(np.float32)
    a[0, :6] = [0.0, -0.0, np.inf, -np.inf, np.nan, 1e-45]
    b = rng.standard_normal((5, 4)) * 1e300
    b[1, 1] = 5e-324
    c = rng.integers(-1000, 1000, (3, 3))
    return [a, b, c]


@pytest.mark.parametrize("k", [0, 1, 2])
def test_csv_bytes_equal_reference_format(tmp_path, k):
    from das_amd.output import save_matrix_csv

    m = _matrices()[k]
    ra = np.arange(m.shape[0]) * 0.005 if k != 2 else np.arange(m.shape[0]) * 1000000
    ca = np.linspace(-3.3, 1e7, m.shape[1]) if k != 1 else np.arange(m.shape[1])
    save_matrix_csv(tmp_path / "a.csv", m, ra, ca, "depth_km", "distance_km")
    OO.save_matrix_csv(tmp_path / "b.csv", m, ra, ca, "depth_km", "distance_km")
    assert (tmp_path / "a.csv").read_bytes() == (tmp_path / "b.csv").read_bytes()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csv_nan_sign_bit(tmp_path, dtype):
    """A NaN with its sign bit set prints "nan", as Python's format does
    (visualizer.py:36-38), in the matrix and in both axes."""
    from das_amd.output import save_matrix_csv

    nneg = np.copysign(np.nan, -1.0)
    m = np.array([[nneg, np.nan, -np.inf], [1.5, nneg, 0.0]], dtype=dtype)
    assert np.signbit(m[0, 0]) and np.signbit(m[1, 1])
    ra = np.array([nneg, 1.0])
    ca = np.array([0.0, nneg, np.nan])
    save_matrix_csv(tmp_path / "a.csv", m, ra, ca, "z", "x")
    OO.save_matrix_csv(tmp_path / "b.csv", m, ra, ca, "z", "x")
    a = (tmp_path / "a.csv").read_bytes()
    assert a == (tmp_path / "b.csv").read_bytes()
    assert b"-nan" not in a
    assert a.splitlines()[1].startswith(b"nan,nan,nan,-inf")


def test_oracle_round_trip(tmp_path):
    m = np.random.default_rng(1).standard_normal((4, 6)).astype(np.float32)
    OO.save_matrix_csv(tmp_path / "o.csv", m, np.arange(4), np.arange(6) * 0.5, "t", "x")
    lines = (tmp_path / "o.csv").read_text().splitlines()
    assert lines[0].startswith("t/x,0,0.5,1,")
    back = np.array([[float(v) for v in ln.split(",")[1:]] for ln in lines[1:]])
    np.testing.assert_allclose(back, m, rtol=1e-6)


def test_axis_mismatch():
    from das_amd.output import save_matrix_csv

    with pytest.raises(AssertionError):
        save_matrix_csv("/dev/null", np.zeros((2, 3)), np.arange(3), np.arange(3))
