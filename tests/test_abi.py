"""The C-ABI library loads and exports every symbol include/few_rocm.h declares (no GPU calls)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "few_rocm.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(few_[A-Za-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = _declared()
    for must in ("few_interpolate_arrays", "few_get_waveform", "few_last_error"):
        assert must in names


def test_library_exports_all_declared_symbols():
    from fastemriwaveforms_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libfew_rocm.so is not built (run python -m fastemriwaveforms_amd._build)")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    assert set(_declared()) == set(_lib.exported_symbols())


def test_error_paths_without_gpu():
    """Argument validation runs on the host: bad sizes return an error code and a
    message instead of exiting (the reference exits the process)."""
    from fastemriwaveforms_amd import _lib
    lib = _lib.lib()
    rc = lib.few_interpolate_arrays(None, None, 0, 0, None, None, None, None, None)
    assert rc != 0
    assert b"null pointer" in lib.few_last_error()
    rc = lib.few_spline_build_async(1, 1, 1, 1, 0, 1, 1, 1, 1, 1, 5, 2, 0, 1, None)
    assert rc != 0 and b"length >= 3" in lib.few_last_error()
    assert lib.few_abi_version() == 1


def test_dropin_precision_setting():
    """The drop-in get_waveform_wrap's precision: FP64 (the reference's) unless set;
    unknown names are refused (no GPU calls)."""
    from fastemriwaveforms_amd import pyinterp
    before = pyinterp.get_precision()
    try:
        pyinterp.set_precision("mixed16")
        assert pyinterp.get_precision() == "mixed16"
        with pytest.raises(ValueError):
            pyinterp.set_precision("fp16")
        assert pyinterp.get_precision() == "mixed16"
    finally:
        pyinterp.set_precision(before)
