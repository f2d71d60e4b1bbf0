"""ctypes bindings for the reference's own table1 C kernels (oracle/_ref).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

``oracle/build_ref.sh`` compiles ``mrrt/nufft/_extensions/c/nufft_table.c``
from the reference tree into ``oracle/_ref/libnufft_table_ref.so``.  This
module restates the Cython bridge around it (``_nufft_table.pyx:29-704``):
planar real/imag copies, a per-repetition loop for the forward
(``pyx:252-254``) and one call with ``N`` repetitions for the adjoint
(OpenMP across repetitions, ``nufft_table.template.c:405``).

Signatures follow ``nufft_table.template.h:34-278``.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libnufft_table_ref.so")
_lib = None


def available():
    return os.path.exists(LIB_PATH)


def _load():
    global _lib
    if _lib is None:
        if not available():
            raise FileNotFoundError(
                "oracle/_ref/libnufft_table_ref.so is not built (oracle/build_ref.sh)")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _prefix(rdt):
    return "float" if np.dtype(rdt) == np.float32 else "double"


def _planar(a, rdt):
    return (np.ascontiguousarray(a.real, dtype=rdt),
            np.ascontiguousarray(a.imag, dtype=rdt))


def _table_args(hs, rdt, cplx):
    arrs, args = [], []
    for h in hs:
        if cplx:
            hr, hi = _planar(h, rdt)
            arrs += [hr, hi]
            args += [_ptr(hr), _ptr(hi)]
        else:
            hr = np.ascontiguousarray(np.real(h), dtype=rdt)
            arrs.append(hr)
            args.append(_ptr(hr))
    return arrs, args


def ref_interp_forward(ck, Kd, hs, Jd, L, tm):
    """Reference ``*_interp{d}_table1_{real,complex}_forward`` per repetition."""
    lib = _load()
    ndim = len(Kd)
    cplx = np.iscomplexobj(hs[0])
    rdt = tm.dtype
    ck = np.asarray(ck)
    if ck.ndim == 1:
        ck = ck[:, None]
    M = tm.shape[0]
    tmF = np.asfortranarray(tm.astype(rdt))
    keep, targs = _table_args(hs, rdt, cplx)
    fn = getattr(lib, "%s_interp%d_table1_%s_forward" % (_prefix(rdt), ndim, "complex" if cplx else "real"))
    fn.restype = None
    cdt = np.result_type(rdt, np.complex64)
    out = np.zeros((M, ck.shape[1]), cdt)
    for r in range(ck.shape[1]):
        rc, ic = _planar(ck[:, r], rdt)
        rf = np.zeros(M, rdt)
        fi = np.zeros(M, rdt)
        args = [_ptr(rc), _ptr(ic)] + [ctypes.c_int(int(k)) for k in Kd] + targs
        args += [ctypes.c_int(int(j)) for j in Jd] + [ctypes.c_int(int(L)), _ptr(tmF),
                 ctypes.c_int(M), _ptr(rf), _ptr(fi)]
        fn(*args)
        out[:, r] = rf + 1j * fi
    return out


def ref_interp_adjoint(fm, Kd, hs, Jd, L, tm):
    """Reference ``*_interp{d}_table1_{real,complex}_adj`` with N repetitions."""
    lib = _load()
    ndim = len(Kd)
    cplx = np.iscomplexobj(hs[0])
    rdt = tm.dtype
    fm = np.asarray(fm)
    if fm.ndim == 1:
        fm = fm[:, None]
    M, N = fm.shape
    tmF = np.asfortranarray(tm.astype(rdt))
    keep, targs = _table_args(hs, rdt, cplx)
    fn = getattr(lib, "%s_interp%d_table1_%s_adj" % (_prefix(rdt), ndim, "complex" if cplx else "real"))
    fn.restype = None
    nk = int(np.prod(Kd))
    rc = np.zeros(nk * N, rdt)
    ic = np.zeros(nk * N, rdt)
    rf = np.ascontiguousarray(fm.real.T, dtype=rdt).ravel()
    fi = np.ascontiguousarray(fm.imag.T, dtype=rdt).ravel()
    args = [_ptr(rc), _ptr(ic)] + [ctypes.c_int(int(k)) for k in Kd] + targs
    args += [ctypes.c_int(int(j)) for j in Jd] + [ctypes.c_int(int(L)), _ptr(tmF),
             ctypes.c_int(M), _ptr(rf), _ptr(fi), ctypes.c_int(N)]
    fn(*args)
    cdt = np.result_type(rdt, np.complex64)
    return (rc + 1j * ic).astype(cdt).reshape((N, nk)).T.copy()
