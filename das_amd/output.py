"""Output formats: CSV image dumps byte-compatible with the reference's
``save_matrix_csv`` (tsw_rtm/visualizer.py:19-39), written by the native
library (formatting a 512 x 2048 image takes a fraction of a second instead of
a Python loop over a million values)."""
import numpy as np

from . import _native as N


def save_matrix_csv(path: str, matrix: np.ndarray, row_axis: np.ndarray, col_axis: np.ndarray,
                    row_name: str = "row", col_name: str = "col"):
    """Row 0: "row_name/col_name" and the column axis (%.6g); each further row:
    its row-axis value (%.6g) and the values (%.6e)."""
    m = np.asarray(matrix)
    nr, nc = m.shape
    assert len(row_axis) == nr and len(col_axis) == nc, "Axis length mismatch"
    if m.dtype == np.float32:
        m, prec = np.ascontiguousarray(m), 4
    else:
        m, prec = np.ascontiguousarray(m, dtype=np.float64), 8
    ra = np.ascontiguousarray(np.asarray(row_axis, dtype=np.float64))
    ca = np.ascontiguousarray(np.asarray(col_axis, dtype=np.float64))
    N.check(N.lib().tsw_save_matrix_csv(str(path).encode(), N.ptr(m), prec, nr, nc, N.ptr(ra),
                                        N.ptr(ca), str(row_name).encode(),
                                        str(col_name).encode()))
    print(f"Saved CSV to {path}")
