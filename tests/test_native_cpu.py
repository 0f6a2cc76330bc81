"""CPU-side checks of the C ABI: the engine library loads and exports every
function include/tswrtm.h declares, and the ctypes table matches the header.
No compute call is made (there is no GPU here)."""
import ctypes
import os
import re

import pytest

from das_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "tswrtm.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_header_lists_reference_symbols():
    names = declared_functions()
    for ref in ("acoustic_2d_step", "elastic_2d_step", "correlate_imaging_condition",
                "build_cerjan_sponge"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(N.LIB_PATH):
        pytest.fail(f"engine library not built: {N.LIB_PATH}")
    lib = ctypes.CDLL(N.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(N.SIGNATURES) == declared_functions()


def test_loader_binds_all_signatures():
    L = N.lib()
    for name in N.SIGNATURES:
        assert getattr(L, name).argtypes is not None
    assert L.tsw_version() == 1


def test_recv_map_error_does_not_need_gpu():
    # argument validation happens before any device call
    import numpy as np

    from das_amd.wave_solver import _norm_indices, _norm_index

    assert list(_norm_indices(np.array([-1, 0, 3]), 10, "rix")) == [9, 0, 3]
    with pytest.raises(IndexError):
        _norm_indices(np.array([10]), 10, "rix")
    with pytest.raises(IndexError):
        _norm_index(-11, 10, "six")
