"""ctypes binding of libpxff.so (the gfx950 C ABI declared in include/pxff.h).

The library is built in-tree by ``pyxtal_ff_amd/csrc/build.sh`` (driven by
``__graft_entry__.build()``).  There is no CPU fallback: if the library is
missing, or no GPU is visible, every entry point raises.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libpxff.so")
_lib = None

P = ctypes.c_void_p
I64 = ctypes.c_int64
I32 = ctypes.c_int
F64 = ctypes.c_double


class PxffError(RuntimeError):
    pass


class So4Params(ctypes.Structure):
    _fields_ = [("twol", ctypes.c_int), ("rcut", ctypes.c_double), ("rfac0", ctypes.c_double),
                ("cutoff_id", ctypes.c_int), ("normalize_u", ctypes.c_int)]


class So3Params(ctypes.Structure):
    _fields_ = [("nmax", ctypes.c_int), ("lmax", ctypes.c_int), ("rcut", ctypes.c_double),
                ("alpha", ctypes.c_double), ("cutoff_id", ctypes.c_int), ("nquad", ctypes.c_int),
                ("weight_on", ctypes.c_int)]


_SIGS = {
    "pxff_last_error": ([ctypes.c_char_p, ctypes.c_size_t], I32),
    "pxff_version": ([], ctypes.c_char_p),
    "pxff_scan_i64": ([P, I64, P, P], I32),
    "pxff_scan_i64_capped": ([P, I64, I64, P, P, P], I32),
    "pxff_tune_set": ([ctypes.c_int, ctypes.c_int], I32),
    "pxff_tune_get": ([ctypes.c_int], I32),
    "pxff_workspace_size": ([ctypes.c_int, P, I64, I64, P], I32),
    "pxff_nlist_count": ([P, I64, P, P, P, P, P, P, I64, F64, P, P, P, P, P, P, P], I32),
    "pxff_nlist_fill": ([P, I64, P, P, P, P, P, F64, P, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_nlist_count_slots": ([P, I64, P, P, P, P, P, P, I64, F64, P, P, P, P, P, P, ctypes.c_int, P, P, P, P, P],
                               I32),
    "pxff_nlist_fill_slots": ([P, I64, P, P, P, P, P, F64, P, P, P, P, P, P, ctypes.c_int, P, P, P, P, P, P, P, P, P],
                              I32),
    "pxff_nlist_rev": ([I64, I64, P, P, P, P, P, P, P], I32),
    "pxff_nlist_rows": ([I64, I64, I64, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_sizes": ([ctypes.c_int, P, P, P, P], I32),
    "pxff_so4_ui": ([P, I64, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_ui_staged": ([P, I64, P, P, P, P, P, P, I64, P, P, P], I32),
    "pxff_so4_zb": ([P, I64, P, P, P, P], I32),
    "pxff_so4_zlen": ([ctypes.c_int, I64, P], I32),
    "pxff_so4_z_atom_major": ([P, I64, P, P, P], I32),
    "pxff_so4_b": ([P, I64, P, P, P, P], I32),
    "pxff_segment_sum": ([I64, ctypes.c_int, P, P, P, P], I32),
    "pxff_mlp_energy_grad": ([I64, ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_yi": ([P, I64, P, P, P, P], I32),
    "pxff_so4_deidrj": ([P, I64, P, P, P, P, P, P, P], I32),
    "pxff_so4_deidrj_counted": ([P, I64, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_gather": ([I64, P, P, P, P, P, P, P], I32),
    "pxff_so4_forces": ([I64, P, P, P, P, P, P, P], I32),
    "pxff_so4_forces_energy": ([I64, P, P, P, P, P, P, P, I64, P, P], I32),
    "pxff_so4_tangent": ([P, I64, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_tangent_noalloc": ([P, I64, P, P, P, P, P, P, P, P], I32),
    "pxff_so4_dbdr": ([P, I64, P, P, P, P, P, P, P], I32),
    "pxff_so4_dbdr_centres": ([P, I64, P, P, P, P, P, I32, P, P], I32),
    "pxff_so4_rows": ([I64, ctypes.c_int, P, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_dxdr_contract_t": ([I64, I32, P, P, P, P, P, P], I32),
    "pxff_dxdr_contract": ([I64, ctypes.c_int, P, P, P, P, P, P, P], I32),
    "pxff_rdxdr_reduce": ([I64, ctypes.c_int, P, P, P, P], I32),
    "pxff_so3_ctot": ([P, I64, P, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so3_ctot_pairs": ([P, I64, P, P, P, P, P, P, P, P, P, P, P, I64, P, P], I32),
    "pxff_so3_radial_table_size": ([P, P, P], I32),
    "pxff_so3_radial_table": ([P, P, P, P, P, P, P], I32),
    "pxff_so3_power": ([P, I64, P, P, P, P], I32),
    "pxff_so3_adjoint_a": ([P, I64, P, P, P, P], I32),
    "pxff_so3_deidrj": ([P, I64, P, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so3_deidrj_counted": ([P, I64, P, P, P, P, P, P, P, P, P, P, P, P], I32),
    "pxff_so3_dpdr": ([P, I64, P, P, P, P, P, P, P, P, P, P, P], I32),
}

OPTIONAL = set()


def exported_symbols():
    """Names this binding requires from the library."""
    return sorted(_SIGS)


def load(require_gpu=True):
    """Load libpxff.so; raise if it is missing (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise PxffError("libpxff.so not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        if not hasattr(L, name):
            if name in OPTIONAL:
                continue
            raise PxffError("libpxff.so lacks symbol %s" % name)
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


_gpu_seen = False


def lib():
    global _gpu_seen
    L = load()
    if not _gpu_seen:  # checked until a GPU has been seen once (it does not go away)
        if not torch.cuda.is_available():
            raise PxffError("pyxtal_ff_amd needs a visible gfx950 GPU (torch.cuda.is_available() is False)")
        _gpu_seen = True
    return L


def last_error():
    buf = ctypes.create_string_buffer(1024)
    load().pxff_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise PxffError("%s failed (%d): %s" % (what, rc, last_error()))


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise PxffError("expected a device tensor")
    if not t.is_contiguous():
        raise PxffError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def stream():
    """torch's current stream on the current device (the raw handle; the
    same value as torch.cuda.current_stream().cuda_stream, fetched without
    the Python stream object)."""
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(torch._C._cuda_getDevice()))


def call(name, *args):
    rc = getattr(lib(), name)(*args)
    check(rc, name)
