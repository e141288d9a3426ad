"""The C-ABI library loads (no GPU needed) and exports every entry point
include/pxff.h declares; the ctypes binding covers them all."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "pxff.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(pxff_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "pxff_so4_deidrj" in names and "pxff_nlist_count" in names
    assert len(names) >= 15


def test_library_exports_declared_symbols():
    from pyxtal_ff_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libpxff.so is not built; run __graft_entry__.build()")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    from pyxtal_ff_amd import _native
    assert set(declared_symbols()) <= set(_native.exported_symbols()) | _native.OPTIONAL


def test_no_gpu_raises_loudly():
    """Without a GPU the product path raises instead of falling back."""
    import torch
    from pyxtal_ff_amd import _native
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_native.PxffError):
        _native.lib()


def test_error_reporting_without_gpu():
    from pyxtal_ff_amd import _native
    L = _native.load()
    v = [ctypes.c_int() for _ in range(4)]
    assert L.pxff_so4_sizes(7, *[ctypes.byref(x) for x in v]) != 0  # odd 2J rejected
    assert "2J" in _native.last_error()
    assert L.pxff_so4_sizes(8, *[ctypes.byref(x) for x in v]) == 0
    assert [x.value for x in v] == [285, 155, 2386, 55]
