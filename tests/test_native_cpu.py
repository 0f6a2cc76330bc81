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


def test_kernel_table_names_and_shapes():
    """The blocked-kernel table (no device call): each configuration's kernel
    family and template arguments agree with its steps, core tile and
    threads; k_tc shapes are 64 - 4T columns by 16R - 4T rows."""
    import re

    from das_amd.batch import kernel_configs, kernel_name

    cfgs = kernel_configs()
    fams = {}
    for kid, (T, W, H, threads) in cfgs.items():
        name = kernel_name(kid)
        m = re.fullmatch(r"(k_t[bsc])<([0-9,]+)>", name)
        assert m, name
        args = [int(v) for v in m.group(2).split(",")]
        fams.setdefault(m.group(1), []).append(kid)
        assert args[0] == T, (kid, name)
        if m.group(1) == "k_tc":
            R = args[1]
            assert (W, H, threads) == (64 - 4 * T, 16 * R - 4 * T, 1024), (kid, name)
        elif m.group(1) == "k_tb":
            _, R, WX, WZ = args
            assert (W, H, threads) == (64 * WX - 4 * T, WZ * R - 4 * T, 64 * WX * WZ), (kid, name)
        else:
            _, WX, Hs = args
            assert (W, H, threads) == (64 * WX - 4 * T, Hs, 64 * WX * (T + 1)), (kid, name)
    assert fams["k_tb"] == list(range(1, 17))
    assert fams["k_ts"] == list(range(17, 25))
    assert fams["k_tc"] == list(range(25, 29))
    L = N.lib()
    buf = ctypes.create_string_buffer(8)
    assert L.tsw_kernel_name(0, buf, 8) == -1
    assert L.tsw_kernel_name(len(cfgs) + 1, buf, 8) == -1
