"""Toeplitz normal operator A^H W A of a table1 NUFFT on one MI355X.

For the NUFFT y_m = sum_n x[n] exp(-i omega_m . (n - s)) the normal operator
is a convolution,

    (A^H W A x)[n] = c * sum_n' x[n'] P[n - n'],   P[d] = sum_m w_m exp(i omega_m . d),

with c the operator's scale (``adjoint_scalefactor``, divided by prod(Kd)
when ``ortho``).  P lives on d in (-N, N); it is computed once by an adjoint
NUFFT of w with image size 2N and n_shift = N (the gridded point-spread
function that ``nufft_adj(..., return_psf=True, grid_only=True)`` exposes in
mrrt/nufft/_nufft.py:1459-1518), embedded in a power-of-two circulant grid
K >= 2N and transformed once.  Each application is then two pruned FFT
passes per axis and one spectrum multiply (fused into the first inverse
pass), with no gather or scatter: ``mi355nufft_toeplitz_*`` in
include/mi355nufft.h.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._nufft import NufftBase, _to_output, _upload_host

__all__ = ["ToeplitzNormal"]


def _fortran_strides(t, ndim):
    """Element strides of axes 0..ndim-1 and of the repetition axis of a
    tensor whose logical shape is Nd + (nreps,)."""
    s = list(t.stride())
    return s[:ndim] + [s[ndim]]


class ToeplitzNormal:
    """``T(x)`` equals ``op.adj(op.fft(x))`` (weighted by ``weights`` when
    given) up to the NUFFT approximation of the point-spread function.

    Parameters
    ----------
    op : NufftBase
        The operator whose normal operator is wanted (its points, precision,
        device, ortho and adjoint_scalefactor are used).
    weights : array or tensor (M,), optional
        Sample weights w_m (e.g. density compensation); default ones.
    psf : tensor, optional
        A precomputed P on the 2N lattice (logical shape 2*Nd, Fortran
        order), already scaled; skips the adjoint NUFFT.
    """

    def __init__(self, op, weights=None, psf=None):
        if not isinstance(op, NufftBase):
            raise TypeError("op must be a NufftBase")
        self.op = op
        self.Nd = tuple(op.Nd)
        self.ndim = len(self.Nd)
        self.device = op.device
        self._tcdt = op._tcdt
        self._lib = _lib.get_lib()
        plan = ctypes.c_void_p()
        _lib.check(self._lib.mi355nufft_toeplitz_create(ctypes.byref(plan), self.ndim, _lib.i64arr(self.Nd),
                                                         0 if op.precision == "single" else 1, None))
        self._plan = plan
        self._stream = None
        kt = (ctypes.c_int64 * self.ndim)()
        _lib.check(self._lib.mi355nufft_toeplitz_grid(self._plan, kt))
        self.Kt = tuple(int(v) for v in kt)
        if psf is None:
            psf = self._psf(weights)
        self.set_psf(psf)

    def _psf(self, weights):
        op = self.op
        N2 = tuple(2 * n for n in self.Nd)
        K2 = tuple(2 * k for k in op.Kd)
        B = NufftBase(N2, op._omega_user, Jd=op.Jd, Kd=K2, Ld=op.Ld, precision=op.precision,
                      n_shift=tuple(float(n) for n in self.Nd), device=self.device)
        if weights is None:
            w = torch.ones((1, op.M), dtype=self._tcdt, device=self.device)
        else:
            w = torch.as_tensor(weights, device=self.device).to(self._tcdt).reshape(1, op.M).contiguous()
        psf = B._adjoint_dev(w, 1)[..., 0]
        c = float(op.adjoint_scalefactor)
        if op.ortho:
            c /= float(np.prod(op.Kd))
        return psf * c

    def set_psf(self, psf):
        """Load P (logical shape 2*Nd, any strides, device tensor)."""
        psf = torch.as_tensor(psf, device=self.device).to(self._tcdt)
        if tuple(psf.shape) != tuple(2 * n for n in self.Nd):
            raise ValueError("psf must have shape 2*Nd = {}".format(tuple(2 * n for n in self.Nd)))
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_toeplitz_set_psf(self._plan, ctypes.c_void_p(psf.data_ptr()),
                                                         _lib.i64arr(list(psf.stride()))))
        self._psf_keep = psf  # the embedding kernel reads it on the stream

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            _lib.check(self._lib.mi355nufft_toeplitz_set_stream(self._plan, ctypes.c_void_p(s)))
            self._stream = s

    def apply_dev(self, x, out=None):
        """x: device tensor with logical shape Nd + (nreps,) (any strides);
        returns A^H W A x with the same logical shape, Fortran-ordered, or
        written into ``out`` (same logical shape, any strides)."""
        nreps = x.shape[-1]
        x = x.to(self._tcdt)
        if out is None:
            out = torch.empty((nreps,) + tuple(reversed(self.Nd)), dtype=self._tcdt, device=self.device)
            out = out.permute(*(tuple(range(self.ndim, 0, -1)) + (0,)))
        elif tuple(out.shape) != tuple(x.shape) or out.dtype != self._tcdt:
            raise ValueError("out must have the logical shape and dtype of the result")
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_toeplitz_apply(
            self._plan, ctypes.c_void_p(x.data_ptr()), _lib.i64arr(_fortran_strides(x, self.ndim)),
            ctypes.c_void_p(out.data_ptr()), _lib.i64arr(_fortran_strides(out, self.ndim)), nreps))
        return out

    def __call__(self, x):
        """A^H W A x for x of shape Nd or Nd + (nreps,), numpy or tensor;
        numpy in, numpy out."""
        was_numpy = not isinstance(x, torch.Tensor)
        if was_numpy:  # through the device's pinned staging buffer
            xt = _upload_host(x, self.device)
        else:
            xt = x.to(self.device)
        squeeze = tuple(xt.shape) == self.Nd
        if squeeze:
            xt = xt.unsqueeze(-1)
        if tuple(xt.shape[:-1]) != self.Nd:
            raise ValueError("x must have shape Nd or Nd + (nreps,)")
        y = self.apply_dev(xt)
        if squeeze:
            y = y[..., 0]
        return _to_output(y, was_numpy)

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            try:
                self._lib.mi355nufft_toeplitz_destroy(plan)
            except Exception:
                pass
            self._plan = None
