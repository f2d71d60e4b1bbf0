"""Multi-coil (SENSE) MRI encoding around the MI355X NUFFT.

    E x = [ A (S_c x) ]_c                       (forward, one image per frame)
    E^H W y = sum_c conj(S_c) A^H (W y_c)       (adjoint)
    E^H W E x                                   (normal; optionally Toeplitz)

with coil sensitivities S_c and real density-compensation weights W.  The
reference recommends MRI_Operator of the external mrrt.operators for this
(mrrt/nufft/_nufft.py:3-5); that package is not in the reference tree, so
this class follows the mathematics above rather than its API.

All coils of all frames go through one transform call: the coil images are
formed by one expand kernel, transformed as repetitions (``NufftBase``) or as
frames x coils (``NufftFrames``), and combined by one kernel
(``mi355nufft_coil_expand`` / ``_coil_combine`` / ``_weight_samples``).
Images and maps are held in torch's natural (C) pixel order; the transforms
see permuted views with the matching strides.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._nufft import NufftBase, _to_output, _upload_host
from .batched import NufftFrames

__all__ = ["SenseOperator"]


class SenseOperator:
    """SENSE encoding operator on one MI355X.

    Parameters
    ----------
    op : NufftBase or NufftFrames
        The transform: one set of points (images of shape Nd) or several
        frames (images of shape (F,) + Nd, ``op.ncoils`` equal to the number
        of maps).
    smaps : array or tensor (C,) + Nd
        Coil sensitivities.
    weights : array or tensor (M,), optional
        Real density-compensation weights applied to the samples in the
        adjoint (and inside the normal operator).
    """

    def __init__(self, op, smaps, weights=None):
        if isinstance(op, NufftFrames):
            self.frames = True
            base = op.op
        elif isinstance(op, NufftBase):
            self.frames = False
            base = op
        else:
            raise TypeError("op must be a NufftBase or NufftFrames")
        self.op = op
        self.base = base
        self.Nd = tuple(base.Nd)
        self.ndim = len(self.Nd)
        self.device = base.device
        self._tcdt, self._trdt = base._tcdt, base._trdt
        self._prec = 0 if base.precision == "single" else 1
        self._lib = _lib.get_lib()
        smaps = torch.as_tensor(smaps).to(self.device).to(self._tcdt).contiguous()
        if smaps.dim() != self.ndim + 1 or tuple(smaps.shape[1:]) != self.Nd:
            raise ValueError("smaps must have shape (C,) + Nd = (C,) + {}".format(self.Nd))
        self.smaps = smaps
        self.ncoils = int(smaps.shape[0])
        if self.frames and op.ncoils != self.ncoils:
            raise ValueError("NufftFrames was built for {} coils, smaps has {}".format(op.ncoils, self.ncoils))
        self.nframes = op.nframes if self.frames else 1
        self.M = base.M
        self.npix = int(np.prod(self.Nd))
        if weights is not None:
            w = torch.as_tensor(weights).to(self.device).to(self._trdt).reshape(-1).contiguous()
            if w.numel() != self.M:
                raise ValueError("weights must have M = {} entries".format(self.M))
            weights = w
        self.weights = weights
        self._toeplitz = None

    # ------------------------------------------------------------ helpers
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _image_shape(self):
        return ((self.nframes,) if self.frames else ()) + self.Nd

    def _expand(self, x):
        """x (F,) + Nd or Nd -> coil images (F*C,) + Nd, contiguous."""
        nimg = self.nframes
        out = torch.empty((nimg * self.ncoils,) + self.Nd, dtype=self._tcdt, device=self.device)
        _lib.check(self._lib.mi355nufft_coil_expand(self._prec, ctypes.c_void_p(x.data_ptr()),
                                                    ctypes.c_void_p(self.smaps.data_ptr()), self.npix, nimg,
                                                    self.ncoils, ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def _combine(self, img):
        nimg = self.nframes
        x = torch.empty(self._image_shape(), dtype=self._tcdt, device=self.device)
        _lib.check(self._lib.mi355nufft_coil_combine(self._prec, ctypes.c_void_p(img.data_ptr()),
                                                     ctypes.c_void_p(self.smaps.data_ptr()), self.npix, nimg,
                                                     self.ncoils, ctypes.c_void_p(x.data_ptr()), self._stream()))
        return x

    def _weighted(self, y):
        if self.weights is None:
            return y
        out = torch.empty_like(y)
        _lib.check(self._lib.mi355nufft_weight_samples(self._prec, ctypes.c_void_p(y.data_ptr()),
                                                       ctypes.c_void_p(self.weights.data_ptr()), self.M,
                                                       y.shape[0], ctypes.c_void_p(out.data_ptr()), self._stream()))
        return out

    def _as_rep_view(self, img):
        """(C,) + Nd contiguous -> logical Nd + (C,) view for NufftBase."""
        return img.permute(*(tuple(range(1, self.ndim + 1)) + (0,)))

    def _in(self, x, shape):
        was_numpy = not isinstance(x, torch.Tensor)
        if was_numpy:  # through the device's pinned staging buffer
            t = _upload_host(x, self.device)
        else:
            t = x.to(self.device)
        if tuple(t.shape) != tuple(shape):
            raise ValueError("expected shape {}, got {}".format(tuple(shape), tuple(t.shape)))
        return t.to(self._tcdt).contiguous(), was_numpy

    @staticmethod
    def _out(t, was_numpy):
        return _to_output(t, was_numpy)

    # ------------------------------------------------------------ device API
    def forward_dev(self, x):
        """x: device tensor of the image shape -> (C, M) samples."""
        img = self._expand(x)
        if self.frames:
            return self.op.forward(img.reshape((self.nframes, self.ncoils) + self.Nd))
        return self.base._forward_dev(self._as_rep_view(img), self.ncoils)

    def adjoint_dev(self, y):
        """y: (C, M) contiguous device tensor -> image."""
        yw = self._weighted(y)
        if self.frames:
            img = self.op.adjoint(yw)
        else:
            img = torch.empty((self.ncoils,) + self.Nd, dtype=self._tcdt, device=self.device)
            self.base._adjoint_dev(yw, self.ncoils, out=self._as_rep_view(img))
        return self._combine(img)

    def normal_dev(self, x, toeplitz=False):
        """E^H W E x; ``toeplitz`` uses the Toeplitz normal operator of the
        transform (one set of points only)."""
        if not toeplitz:
            return self.adjoint_dev(self.forward_dev(x))
        if self.frames:
            raise ValueError("the Toeplitz normal operator needs one set of points (NufftBase)")
        if self._toeplitz is None:
            from .toeplitz import ToeplitzNormal
            self._toeplitz = ToeplitzNormal(self.base, weights=self.weights)
        img = self._expand(x)
        out = torch.empty_like(img)
        self._toeplitz.apply_dev(self._as_rep_view(img), out=self._as_rep_view(out))
        return self._combine(out)

    # ------------------------------------------------------------ array API
    def forward(self, x):
        """Image (numpy or tensor, shape Nd or (F,) + Nd) -> (C, M) samples."""
        t, was_numpy = self._in(x, self._image_shape())
        return self._out(self.forward_dev(t), was_numpy)

    def adjoint(self, y):
        """(C, M) samples -> image."""
        t, was_numpy = self._in(y, (self.ncoils, self.M))
        return self._out(self.adjoint_dev(t), was_numpy)

    def normal(self, x, toeplitz=False):
        t, was_numpy = self._in(x, self._image_shape())
        return self._out(self.normal_dev(t, toeplitz), was_numpy)
