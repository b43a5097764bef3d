"""The C ABI library loads and exports every entry point include/pa_hip.h
declares; argument validation needs no device (no compute calls here)."""
import ctypes
import os
import re

import pytest

from pyatmosphere_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "pa_hip.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pa_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_bound_symbols():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_library_exports_every_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_constants_match_header():
    text = open(HEADER).read()
    assert f"#define PA_MAX_PUPILS {_lib.PA_MAX_PUPILS}" in text
    assert "#define PA_CH_PARAMS 4" in text
    assert f"#define PA_LOWRANK_TERMS {_lib.PA_LOWRANK_TERMS}" in text
    assert "#define PA_SCREEN_PARAMS (8 + PA_LOWRANK_TERMS * PA_LOWRANK_TERMS + 2 * PA_LOWRANK_TERMS)" in text
    assert _lib.PA_SCREEN_PARAMS == 8 + _lib.PA_LOWRANK_TERMS ** 2 + 2 * _lib.PA_LOWRANK_TERMS


def test_size_validation_without_device():
    lib = _lib.load()
    assert lib.pa_tile_rows(2048) == 4
    assert lib.pa_tile_rows(1000) == -2
    assert b"power of two" in lib.pa_last_error()
    assert lib.pa_operand_bytes(2048, 1024, 0) == 2 * 2 * 2048 * 2048
    assert lib.pa_operand_bytes(2048, 1024, 2) == 4 * 2048 * 2048
    assert lib.pa_operand_bytes(64, 10, 0) == 2 * 2 * 64 * 32        # K padded to 32
    with pytest.raises(RuntimeError, match="batch"):
        _lib.call("pa_row_first", 256, 1, None, 1, None, 1, 0, None)


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.NativeLibraryError):
        _lib.load()


def test_argument_validation_without_device():
    """Bad modes, launch shapes, sizes and null pointers are rejected before
    anything is launched (fake non-null pointers are never dereferenced)."""
    P = 1
    with pytest.raises(RuntimeError, match="GEMM mode"):
        _lib.call("pa_screen_gemm", 2048, 1024, P, P, P, P, None, None, P, 7, 0, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="tile"):
        _lib.call("pa_screen_gemm", 2048, 1024, P, P, P, P, None, None, P, 0, 3, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_screen_gemm", 2048, 1024, P, P, P, P, None, None, None, 0, 0, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="rows"):
        _lib.call("pa_screen_tables", 2048, 1024, P, P, P, None, P, P, P, P, 0, 5, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="points and batch"):
        _lib.call("pa_screen_tables", 2048, 0, P, P, P, None, P, P, P, P, 0, 0, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="power of two"):
        _lib.call("pa_col_step", 3000, P, P, None, None, 0, 1, P, 0, 1, None)
    with pytest.raises(RuntimeError, match="cols"):
        _lib.call("pa_col_step", 2048, P, P, None, None, 0, 1, P, 3, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_row_mid", 2048, P, None, None, 0, P, 1, None)
    from pyatmosphere_amd import autotune
    assert autotune.col_groups(2048) == (4, 2, 8) and autotune.col_groups(4096) == (2, 1, 4)
    with pytest.raises(RuntimeError, match="count"):
        _lib.call("pa_expmi", -1, None, None, None)


def test_new_entry_points_validate_without_device():
    """pa_screen_draw_prep and pa_structure_function reject bad sizes, modes and
    null pointers before launching; the structure function's scratch is
    (n / 4) row groups x (n - 1) lags of float64 per screen."""
    P = 1
    with pytest.raises(RuntimeError, match="points and batch"):
        _lib.call("pa_screen_draw_prep", 0, P, P, None, P, 1, 0, P, P, 0, P, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_screen_draw_prep", 1024, P, None, None, P, 1, 0, P, P, 0, P, None, 0, 1, None)
    with pytest.raises(RuntimeError, match="GEMM mode"):
        _lib.call("pa_screen_draw_prep", 1024, P, P, None, P, 1, 0, P, P, 7, P, None, 0, 1, None)
    assert _lib.call("pa_structure_function_scratch", 256, 2) == 2 * (256 // 4) * 255 * 8
    assert _lib.load().pa_structure_function_scratch(1000, 1) < 0
    with pytest.raises(RuntimeError, match="power of two"):
        _lib.call("pa_structure_function", 1000, P, P, P, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_structure_function", 256, P, None, P, 1, None)


def test_separable_first_step_entry_points_validate_without_device():
    """pa_source_fresnel and pa_row_mid_source reject bad grid sizes, null
    pointers, batch <= 0 and max_steps < 1 before launching."""
    P = 1
    with pytest.raises(RuntimeError, match="power of two"):
        _lib.call("pa_source_fresnel", 1000, P, P, P, None, 1, P, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_source_fresnel", 2048, None, P, P, None, 1, P, 1, None)
    with pytest.raises(RuntimeError, match="batch"):
        _lib.call("pa_source_fresnel", 2048, P, P, P, None, 1, P, 0, None)
    with pytest.raises(RuntimeError, match="max_steps"):
        _lib.call("pa_source_fresnel", 2048, P, P, P, None, 0, P, 1, None)
    with pytest.raises(RuntimeError, match="power of two"):
        _lib.call("pa_row_mid_source", 100, P, P, P, P, 1, None)
    with pytest.raises(RuntimeError, match="null"):
        _lib.call("pa_row_mid_source", 2048, P, None, P, P, 1, None)
    with pytest.raises(RuntimeError, match="batch"):
        _lib.call("pa_row_mid_source", 2048, P, P, P, P, -1, None)
