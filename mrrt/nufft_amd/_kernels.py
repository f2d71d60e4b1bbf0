"""Separable NUFFT interpolation kernels (the public kernel classes of
mrrt/nufft/_kernels.py:18-162, same names and attributes).

``KaiserBesselKernel(shape, alpha, m)`` holds one 1-D generalized
Kaiser-Bessel function per axis in ``kernels``; ``BeattyKernel(shape,
grid_shape, os_grid_shape)`` is the order-0 KB window with Beatty's shape
parameter (Beatty et al., IEEE TMI 2005, eq. 5), the kernel every table of
this package is built from.
"""

import functools

import numpy as np

from ._kernel_math import beatty_alpha, kaiser_bessel

__all__ = ["NufftKernelBase", "KaiserBesselKernel", "BeattyKernel"]

_KERNEL_DOC = """Interpolation kernel.

Parameters
----------
k : ndarray
    Coordinates at which to evaluate the kernel, within [-J/2, J/2] where J
    is the kernel size on this axis.
"""


def _per_axis_ints(v, n):
    """One integer per axis, as the reference's _as_1d_ints
    (_utils.py:89-127): a scalar is broadcast, non-integer values raise."""
    v = np.atleast_1d(np.squeeze(np.asarray(v)))
    if v.ndim > 1:
        raise ValueError("arr must be scalar or 1d")
    if not issubclass(v.dtype.type, np.integer) and not np.all(np.mod(v, 1) == 0):
        raise ValueError("arr contains non-integer values")
    if v.size != n:
        if v.size != 1:
            raise ValueError("array did not have the expected size of {}".format(n))
        v = np.repeat(v, n)
    return tuple(int(a) for a in v)


class NufftKernelBase(object):
    """Common attributes: ``shape`` (taps per axis), ``kernels`` (one 1-D
    function per axis), ``kernel_type``, ``params``."""

    shape = ()
    kernels = []
    kernel_type = ""
    params = {}

    @property
    def ndim(self):
        return len(self.shape)

    def plot(self, real_imag=False, axes=None):
        """Plot the kernel of each axis over [-J/2, J/2] (needs matplotlib)."""
        from matplotlib import pyplot as plt
        if axes is None:
            _, axes = plt.subplots(self.ndim, 1, sharex=True)
        axes = np.atleast_1d(axes)
        for d in range(self.ndim):
            j = self.shape[d]
            x = np.linspace(-j / 2, j / 2, 1001)
            y = self.kernels[d](x)
            if real_imag:
                axes[d].plot(x, np.real(y), "k-", label="real")
                axes[d].plot(x, np.imag(y), "r--", label="imaginary")
            else:
                axes[d].plot(x, np.abs(y), "k-", label="magnitude")
            axes[d].set_ylabel("kernel amplitude, axis {}".format(d))
            axes[d].set_title("type: {}".format(self.kernel_type))
            axes[d].grid(True)
            axes[d].legend()
        plt.draw()
        return axes

    def __str__(self):
        return "kernel type: {}\nkernel shape: {}\n".format(self.kernel_type, self.shape)


class KaiserBesselKernel(NufftKernelBase):
    """Generalized Kaiser-Bessel kernel with a given shape parameter
    ``alpha`` and order ``m`` per axis (mrrt/nufft/_kernels.py:96-126).  As
    in the reference, both are taken as integers (non-integer values
    raise); BeattyKernel sets its non-integer alphas directly."""

    is_kaiser_scale = None

    def __init__(self, shape, alpha=None, m=None):
        self.kernel_type = "kb:user"
        self.params = {}
        self.shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
        if alpha is None or m is None:
            raise ValueError("kwargs must contain shape, m, alpha for {} case".format(self.kernel_type))
        self.kernels = self._initialize_kernel(_per_axis_ints(alpha, self.ndim), _per_axis_ints(m, self.ndim))

    def _initialize_kernel(self, alpha, m):
        self.is_kaiser_scale = True
        self.alpha = list(alpha)
        self.m = list(m)
        kernels = []
        for j, a, mm in zip(self.shape, alpha, m):
            k = functools.partial(kaiser_bessel, J=j, alpha=a, m=mm)
            k.__doc__ = _KERNEL_DOC
            kernels.append(k)
        return kernels


class BeattyKernel(KaiserBesselKernel):
    """Order-0 Kaiser-Bessel kernel with the shape parameter of Beatty et
    al. for a grid of ``grid_shape`` oversampled to ``os_grid_shape``
    (mrrt/nufft/_kernels.py:129-162)."""

    def __init__(self, shape, grid_shape, os_grid_shape):
        self.kernel_type = "kb:beatty"
        self.params = {}
        self.shape = (int(shape),) if np.isscalar(shape) else tuple(int(s) for s in shape)
        self.grid_shape = tuple(int(v) for v in np.broadcast_to(np.atleast_1d(grid_shape), (self.ndim,)))
        self.os_grid_shape = tuple(int(v) for v in np.broadcast_to(np.atleast_1d(os_grid_shape), (self.ndim,)))
        alpha, m = self._get_beatty_kb_params()
        self.kernels = self._initialize_kernel(alpha, m)

    def _get_beatty_kb_params(self):
        self.is_kaiser_scale = True
        alpha = [beatty_alpha(j, k, n) for j, k, n in zip(self.shape, self.os_grid_shape, self.grid_shape)]
        return alpha, [0] * self.ndim
