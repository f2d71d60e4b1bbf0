"""ctypes binding of ``libmi355nufft.so`` (declarations: include/mi355nufft.h).

torch is imported before the library is opened so that the library binds to
the HIP runtime and rocFFT that torch has already loaded (same SONAMEs
``libamdhip64.so.7`` / ``librocfft.so.0``); one process then has exactly one
HIP runtime and device pointers from torch are valid in the library.

There is no CPU fallback: if the library is missing or cannot be loaded,
``get_lib()`` raises.
"""

import ctypes
import os
import threading

import torch  # noqa: F401  (must precede the CDLL load, see module docstring)

from .build import LIB_PATH

__all__ = ["get_lib", "check", "NufftLibraryError", "EXPORTED_SYMBOLS"]

_lock = threading.Lock()
_lib = None

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_dblp = ctypes.POINTER(ctypes.c_double)
c_vpp = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); mirrors include/mi355nufft.h
_SIGNATURES = {
    "mi355nufft_last_error": (ctypes.c_char_p, []),
    "mi355nufft_abi_version": (c_int, []),
    "mi355nufft_plan_create": (c_int, [ctypes.POINTER(c_vp), c_int, c_i64p, c_i64p, c_i32p, c_i32, c_int, c_int,
                                       c_int, c_dbl, c_vpp, ctypes.POINTER(c_dblp), c_dblp, c_vp]),
    "mi355nufft_plan_destroy": (c_int, [c_vp]),
    "mi355nufft_set_stream": (c_int, [c_vp, c_vp]),
    "mi355nufft_set_points": (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_dblp, c_int, c_int]),
    "mi355nufft_set_points_frames": (c_int, [c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_dblp, c_int, c_int, c_i64p,
                                             c_i32]),
    "mi355nufft_num_points": (c_i64, [c_vp]),
    "mi355nufft_set_binning": (c_int, [c_vp, c_i32p, c_i32, c_int]),
    "mi355nufft_set_tiling": (c_int, [c_vp, c_i32p, c_i32, c_int, c_i32]),
    "mi355nufft_set_timing": (c_int, [c_vp, c_i32]),
    "mi355nufft_read_timing": (c_int, [c_vp, c_i32, c_i64p, c_dblp]),
    "mi355nufft_read_flush_atomics": (c_int, [c_vp, ctypes.POINTER(ctypes.c_uint64)]),
    "mi355nufft_plan_info": (c_int, [c_vp, c_i64p, c_i32]),
    "mi355nufft_forward": (c_int, [c_vp, c_vp, c_i64p, c_vp, c_i32, c_vp]),
    "mi355nufft_adjoint": (c_int, [c_vp, c_vp, c_vp, c_i64p, c_i32, c_vp]),
    "mi355nufft_deapod_pad": (c_int, [c_vp, c_vp, c_i64p, c_vp, c_i32]),
    "mi355nufft_fft": (c_int, [c_vp, c_vp, c_i32, c_int]),
    "mi355nufft_interp": (c_int, [c_vp, c_vp, c_vp, c_i32]),
    "mi355nufft_spread": (c_int, [c_vp, c_vp, c_vp, c_i32]),
    "mi355nufft_interp_grad": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i32]),
    "mi355nufft_crop_deapod": (c_int, [c_vp, c_vp, c_vp, c_i64p, c_i32]),
    "mi355nufft_forward_grid": (c_int, [c_vp, c_vp, c_i64p, c_vp, c_i32]),
    "mi355nufft_adjoint_image": (c_int, [c_vp, c_vp, c_vp, c_i64p, c_i32]),
    "mi355nufft_set_fft": (c_int, [c_vp, c_int]),
    "mi355nufft_set_scatter": (c_int, [c_vp, c_int]),
    "mi355nufft_measure_atomics": (c_int, [c_int, c_dblp, c_vp]),
    "mi355nufft_set_sparse": (c_int, [c_vp, c_vp, c_vp, c_vp]),
    "mi355nufft_clear_sparse": (c_int, [c_vp]),
    "mi355nufft_coil_expand": (c_int, [c_int, c_vp, c_vp, c_i64, c_i32, c_i32, c_vp, c_vp]),
    "mi355nufft_coil_combine": (c_int, [c_int, c_vp, c_vp, c_i64, c_i32, c_i32, c_vp, c_vp]),
    "mi355nufft_weight_samples": (c_int, [c_int, c_vp, c_vp, c_i64, c_i32, c_vp, c_vp]),
    "mi355nufft_sparse_apply": (c_int, [c_vp, c_int, c_vp, c_vp, c_i32]),
    "mi355nufft_toeplitz_create": (c_int, [ctypes.POINTER(c_vp), c_int, c_i64p, c_int, c_vp]),
    "mi355nufft_toeplitz_destroy": (c_int, [c_vp]),
    "mi355nufft_toeplitz_set_stream": (c_int, [c_vp, c_vp]),
    "mi355nufft_toeplitz_grid": (c_int, [c_vp, c_i64p]),
    "mi355nufft_toeplitz_set_psf": (c_int, [c_vp, c_vp, c_i64p]),
    "mi355nufft_toeplitz_apply": (c_int, [c_vp, c_vp, c_i64p, c_vp, c_i64p, c_i32]),
    "mi355nufft_table1_forward": (c_int, [c_int, c_i64p, c_i32p, c_i32, c_int, c_int, c_vp, c_vpp, c_vp, c_i64,
                                          c_vp, c_i32, c_vp]),
    "mi355nufft_table1_adjoint": (c_int, [c_int, c_i64p, c_i32p, c_i32, c_int, c_int, c_vp, c_vpp, c_vp, c_i64,
                                          c_vp, c_i32, c_vp]),
    "mi355nufft_table1_adjoint_noalloc": (c_int, [c_int, c_i64p, c_i32p, c_i32, c_int, c_int, c_vp, c_vpp, c_vp,
                                                  c_i64, c_vp, c_i32, c_vp]),
}

EXPORTED_SYMBOLS = tuple(_SIGNATURES)


class NufftLibraryError(RuntimeError):
    """Raised when the native library fails or cannot be loaded."""


def get_lib():
    """Load (once) and return the ctypes handle of libmi355nufft.so."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NufftLibraryError(
                    "%s is not built; run `python -m mrrt.nufft_amd.build` "
                    "(there is no CPU fallback)" % LIB_PATH)
            lib = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc):
    """Raise NufftLibraryError with the library's message if rc != 0."""
    if rc != 0:
        msg = get_lib().mi355nufft_last_error()
        msg = msg.decode() if msg else ""
        if rc == 1:
            raise ValueError("mi355nufft: " + msg)
        raise NufftLibraryError("mi355nufft error %d: %s" % (rc, msg))


def i64arr(vals):
    return (ctypes.c_int64 * len(vals))(*[int(v) for v in vals])


def i32arr(vals):
    return (ctypes.c_int32 * len(vals))(*[int(v) for v in vals])


def dblarr(vals):
    return (ctypes.c_double * len(vals))(*[float(v) for v in vals])
