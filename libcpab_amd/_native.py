"""ctypes binding of libcpab_hip.so (the C ABI declared in include/cpab_hip.h).

The library is built in-tree by libcpab_amd/build.py (``hipcc`` for gfx950)
and loaded from libcpab_amd/_lib/.  There is no fallback: a missing or
unloadable library raises, so a GPU call can never silently run elsewhere.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libcpab_hip.so")

_c_int = ctypes.c_int
_c_float = ctypes.c_float
_c_size = ctypes.c_size_t
_c_ptr = ctypes.c_void_p


class CpabTess(ctypes.Structure):
    """Mirror of cpab_tess_t."""
    _fields_ = [("ndim", _c_int), ("nc", _c_int * 3), ("nC", _c_int), ("d", _c_int),
                ("B", _c_ptr), ("planes", _c_ptr), ("nbr", _c_ptr), ("unfinished", _c_ptr), ("nV", _c_int),
                ("amap", _c_ptr)]


class CpabInstance(ctypes.Structure):
    """Mirror of cpab_instance_t."""
    _fields_ = [("tess", CpabTess), ("C", _c_int), ("dsize", _c_int * 3), ("outsize", _c_int * 3), ("nP", _c_int),
                ("item0", _c_int), ("tiles", _c_int), ("theta", _c_ptr), ("data", _c_ptr), ("lin", _c_ptr),
                ("out", _c_ptr), ("saved", _c_ptr), ("grad_out", _c_ptr), ("grad_data", _c_ptr),
                ("grad_theta", _c_ptr), ("ws", _c_ptr), ("ws_bytes", _c_size)]


# name -> (restype, argtypes)
_SIGNATURES = {
    "cpab_status_string": (ctypes.c_char_p, [_c_int]),
    "cpab_abi_version": (_c_int, []),
    "cpab_workspace_bytes": (_c_size, [ctypes.POINTER(CpabTess), _c_int, _c_int]),
    "cpab_data_workspace_bytes": (_c_size, [ctypes.POINTER(CpabTess), _c_int, _c_ptr]),
    "cpab_saved_floats": (_c_size, [_c_int, _c_int, _c_int]),
    "cpab_jac_floats": (_c_size, [_c_int, _c_int, _c_int]),
    "cpab_theta_to_A": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_ptr]),
    "cpab_transform_fwd": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_ptr,
                                    _c_float, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_transform_bwd": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_ptr,
                                    _c_float, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_transform_cells": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_float,
                                      _c_int, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_interpolate_fwd": (_c_int, [_c_int, _c_ptr, _c_int, _c_int, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr]),
    "cpab_interpolate_bwd": (_c_int, [_c_int, _c_ptr, _c_int, _c_int, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                      _c_ptr, _c_ptr]),
    "cpab_transform_data_fwd": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_ptr, _c_ptr,
                                         _c_ptr, _c_float, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_transform_data_fwd_noalloc": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_ptr,
                                                 _c_ptr, _c_ptr, _c_float, _c_ptr, _c_ptr, _c_ptr]),
    "cpab_transform_data_bwd": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_ptr, _c_ptr,
                                         _c_ptr, _c_float, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_findcell": (_c_int, [_c_int, _c_ptr, _c_ptr, _c_int, _c_ptr, _c_ptr]),
    "cpab_vectorfield": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_ptr, _c_int, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_transform_data_fwd_plain": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_ptr,
                                               _c_ptr, _c_ptr, _c_float, _c_ptr, _c_ptr, _c_ptr, _c_size, _c_ptr]),
    "cpab_transform_data_bwd_plain": (_c_int, [ctypes.POINTER(CpabTess), _c_ptr, _c_int, _c_ptr, _c_int, _c_ptr,
                                               _c_ptr, _c_ptr, _c_float, _c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_ptr,
                                               _c_size, _c_ptr]),
    "cpab_batch_workspace_bytes": (_c_size, [ctypes.POINTER(CpabTess), _c_ptr, _c_int]),
    "cpab_batch_plan": (_c_int, [_c_ptr, _c_int, _c_int, _c_ptr, _c_int]),
    "cpab_batch_fwd": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_float, _c_ptr, _c_ptr]),
    "cpab_batch_bwd": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_float, _c_ptr, _c_ptr]),
    "cpab_batch_fwd_plain": (_c_int, [_c_ptr, _c_int, _c_float, _c_ptr]),
    "cpab_batch_bwd_plain": (_c_int, [_c_ptr, _c_int, _c_float, _c_ptr]),
    "cpab_fill_zero": (_c_int, [_c_ptr, _c_size, _c_ptr]),
    "cpab_expm_affine": (_c_int, [_c_ptr, _c_int, _c_int, _c_float, _c_ptr, _c_ptr]),
    "cpab_ref_forward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_int, _c_ptr]),
    "cpab_slow_forward": (_c_int, [_c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_int, _c_ptr, _c_ptr]),
    "cpab_slow_backward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_int, _c_ptr,
                                    _c_ptr, _c_ptr, _c_ptr]),
    "cpab_ref_backward": (_c_int, [_c_ptr, _c_ptr, _c_ptr, _c_ptr, _c_int, _c_ptr, _c_int, _c_int, _c_int, _c_int,
                                   _c_int, _c_int, _c_ptr]),
}

EXPORTED = tuple(_SIGNATURES)

# CPAB_ABI_VERSION of include/cpab_hip.h that these signatures and structs mirror
ABI_VERSION = 7

# static scratch per GPU of cpab_transform_data_fwd_noalloc (CPAB_NOALLOC_SCRATCH_BYTES)
CPAB_NOALLOC_SCRATCH_BYTES = 4 << 20

_lib = None


class CpabError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CpabError("libcpab_hip.so is not built (expected at %s); run "
                            "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.cpab_abi_version() != ABI_VERSION:
            raise CpabError("libcpab_hip.so has ABI version %d, this binding expects %d: rebuild it"
                            % (L.cpab_abi_version(), ABI_VERSION))
        _lib = L
    return _lib


def check(status):
    if status != 0:
        msg = lib().cpab_status_string(status).decode()
        raise CpabError("libcpab_hip: %s (status %d)" % (msg, status))


def int3(vals):
    arr = (_c_int * 3)()
    for i, v in enumerate(list(vals)[:3]):
        arr[i] = int(v)
    for i in range(len(vals), 3):
        arr[i] = 1
    return arr
