"""Batched coils x frames NUFFT with per-frame trajectories.

``NufftFrames`` holds one plan for F frames, each frame with its own
sampling points (uneven sample counts M_f) and C coil images.  A transform
is three launches plus rocFFT for all F*C images at once:

* deapodization + zero-pad of all F*C images (one kernel),
* one batched rocFFT of the F*C oversampled grids,
* one persistent tiled gather (or scatter) over every (frame, subproblem)
  work item, each coil reusing the per-point coefficients.

``capture()`` records the forward and adjoint into HIP graphs (torch's
CUDAGraph API on ROCm) so that a transform is a single graph launch.  Its
eager warm-up calls leave the plan's adjoint grid all zero, so the captured
adjoint has no grid-clearing pass; every replay leaves the grid zero again.
The reference has no batched path: its repetitions are a Python loop of
one kernel launch per repetition (mrrt/nufft/_nufft.py:1069-1084, :1181-1189).
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._nufft import NufftBase

__all__ = ["NufftFrames"]


class NufftFrames(object):
    """Forward/adjoint NUFFT of F frames x C coils.

    Parameters
    ----------
    Nd, Jd, Kd, Ld, precision, phasing, n_shift, ortho : as ``NufftBase``
    omegas : list of F arrays (M_f, ndim), the points of each frame
    ncoils : C, images per frame
    bin_size, max_points : binning of the tiled kernels (None / 0: defaults,
        or the values chosen by ``autotune.tune``)
    """

    def __init__(self, Nd, omegas, ncoils, Jd=6, Kd=None, Ld=1024, precision="single", phasing="real",
                 n_shift=None, ortho=False, device=None, bin_size=None, max_points=0):
        omegas = [np.asarray(o, dtype=np.float64) for o in omegas]
        self.nframes = len(omegas)
        self.ncoils = int(ncoils)
        self.frame_offsets = np.concatenate([[0], np.cumsum([o.shape[0] for o in omegas])]).astype(np.int64)
        om = np.concatenate(omegas, axis=0)
        self.op = NufftBase(Nd, om, Jd=Jd, Kd=Kd, Ld=Ld, precision=precision, phasing=phasing, n_shift=n_shift,
                            ortho=ortho, device=device)
        if bin_size is not None or max_points:
            self.op.set_binning(bin_size, max_points, "auto")
        self.op.set_points(om, frame_offsets=self.frame_offsets)
        self.Nd = self.op.Nd
        self.M = self.op.M
        self.device = self.op.device
        nimg = self.nframes * self.ncoils
        self._grid = torch.empty((nimg, int(np.prod(self.op.Kd))), dtype=self.op._tcdt, device=self.device)
        self._graphs = None

    def frame_slice(self, f):
        """Slice of frame f's samples in the concatenated (C, M) layout."""
        return slice(int(self.frame_offsets[f]), int(self.frame_offsets[f + 1]))

    def _image_view(self, x):
        F, C = self.nframes, self.ncoils
        if tuple(x.shape) != (F, C) + tuple(self.Nd):
            raise ValueError("expected images of shape %s" % (((F, C) + tuple(self.Nd)),))
        xv = x.reshape((F * C,) + tuple(self.Nd))
        return xv.permute(*(tuple(range(1, self.op.ndim + 1)) + (0,)))

    def forward(self, x, out=None):
        """x: (F, C, *Nd) complex device tensor -> y: (C, M) samples."""
        op = self.op
        xv = self._image_view(x)
        y = out if out is not None else torch.empty((self.ncoils, self.M), dtype=op._tcdt, device=self.device)
        op._sync_stream()
        _lib.check(op._lib.mi355nufft_forward(op._plan, ctypes.c_void_p(xv.data_ptr()),
                                              _lib.i64arr(list(xv.stride())), ctypes.c_void_p(y.data_ptr()),
                                              self.ncoils, ctypes.c_void_p(self._grid.data_ptr())))
        return y

    def adjoint(self, y, out=None):
        """y: (C, M) complex device tensor -> x: (F, C, *Nd) images."""
        op = self.op
        if tuple(y.shape) != (self.ncoils, self.M) or not y.is_contiguous():
            raise ValueError("expected contiguous samples of shape (%d, %d)" % (self.ncoils, self.M))
        x = out if out is not None else torch.empty((self.nframes, self.ncoils) + tuple(self.Nd), dtype=op._tcdt,
                                                    device=self.device)
        xv = self._image_view(x)
        op._sync_stream()
        _lib.check(op._lib.mi355nufft_adjoint(op._plan, ctypes.c_void_p(y.data_ptr()),
                                              ctypes.c_void_p(xv.data_ptr()), _lib.i64arr(list(xv.stride())),
                                              self.ncoils, None))  # plan-owned, kept zero
        return x

    def capture(self, x, y, xa):
        """Record forward (x -> y) and adjoint (y -> xa) into two HIP graphs.

        The buffers stay bound to the graphs; ``replay_forward`` /
        ``replay_adjoint`` rerun them on the current contents.
        """
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            for _ in range(2):  # creates the rocFFT plans and work buffers outside capture
                self.forward(x, out=y)
                self.adjoint(y, out=xa)
        torch.cuda.current_stream(self.device).wait_stream(s)
        torch.cuda.synchronize(self.device)
        gf = torch.cuda.CUDAGraph()
        with torch.cuda.graph(gf):
            self.forward(x, out=y)
        ga = torch.cuda.CUDAGraph()
        with torch.cuda.graph(ga):
            self.adjoint(y, out=xa)
        self.op._sync_stream()
        self._graphs = (gf, ga)
        return self._graphs

    def replay_forward(self):
        self._graphs[0].replay()

    def replay_adjoint(self):
        self._graphs[1].replay()
