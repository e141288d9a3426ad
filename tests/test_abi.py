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


def test_workspace_sizes_without_gpu():
    """pxff_workspace_size (a host-side query) against the buffer sizes the
    entry points document."""
    from pyxtal_ff_amd import _native
    L = _native.load()
    b = ctypes.c_int64()
    n = 4097
    p4 = _native.So4Params(8, 4.9, 0.99363, 1, 0)
    assert L.pxff_workspace_size(1, ctypes.byref(p4), n, 0, ctypes.byref(b)) == 0
    zl = ctypes.c_int64()
    assert L.pxff_so4_zlen(8, n, ctypes.byref(zl)) == 0
    assert b.value == 8 * zl.value
    # tangent: complex half rows (155 at 2J=8) of 65 padded tiles of 64 atoms
    assert L.pxff_workspace_size(2, ctypes.byref(p4), n, 0, ctypes.byref(b)) == 0
    assert b.value == 65 * 64 * 155 * 16
    # neighbour-list int32 workspace
    assert L.pxff_workspace_size(0, None, n, 1000, ctypes.byref(b)) == 0
    assert b.value == 4 * (1000 + 1001 + n + 3 * n + n)
    # SO3 pair ordering: 256 x 64 bin counts, the count, the capacity's order
    assert L.pxff_workspace_size(3, None, 0, 576000, ctypes.byref(b)) == 0
    assert b.value == 8 * (256 * 64 + 1 + 576000)
    # SO3 radial table: intervals = the least power of two >= 256 covering
    # 2.5 x the steepness 2 alpha rc^2 (at most 4096), 12 coefficients for
    # each of (lmax + 1) 2 nmax functions
    for nmax, lmax, rc, alpha in [(6, 6, 5.0, 2.0), (3, 4, 5.0, 6.0), (2, 1, 4.0, 0.5)]:
        p3 = _native.So3Params(nmax, lmax, rc, alpha, 1, 10 * (nmax + lmax + 1), 0)
        assert L.pxff_workspace_size(4, ctypes.byref(p3), 0, 0, ctypes.byref(b)) == 0
        nk = 256
        while nk < 4096 and nk < 2.5 * 2 * alpha * rc * rc:
            nk *= 2
        assert b.value == 8 * nk * (lmax + 1) * 2 * nmax * 12, (nmax, lmax, rc, alpha)
    # neighbour-list slots: j int64, T 3 x int32, r 3 x double per entry
    assert L.pxff_workspace_size(6, None, n, 64, ctypes.byref(b)) == 0
    assert b.value == n * 64 * (8 + 3 * 4 + 3 * 8)
    # staged U_tot records: 6 doubles per image of the capacity
    assert L.pxff_workspace_size(5, None, 0, 1000, ctypes.byref(b)) == 0
    assert b.value == 1000 * 6 * 8
    assert L.pxff_workspace_size(9, None, 0, 0, ctypes.byref(b)) != 0
    assert "workspace kind" in _native.last_error()
