"""Exact non-uniform discrete Fourier transform on the GPU (the reference's
test-time ``dtft`` / ``dtft_adj``, mrrt/nufft/_dtft.py:16-216, same
signatures and shape rules).

    X[m] = sum_n x[n] exp(-i omega_m . (n - n_shift))     (dtft)
    x[n] = sum_m X[m] exp(+i omega_m . (n - n_shift))     (dtft_adj)

with n over the Fortran-ordered grid of ``shape``.  Evaluated in float64 /
complex128 on the device as chunked matrix products (chunks of samples,
so the phase block stays under about 64 M entries; ``useloop`` only lowers
that bound).  numpy inputs give numpy results, device tensors give device
tensors.  O(M prod(shape)) work: a reference for testing, not a transform.
"""

import numpy as np
import torch

__all__ = ["dtft", "dtft_adj"]


def _device(obj, device):
    if device is not None:
        return torch.device(device)
    if isinstance(obj, torch.Tensor) and obj.is_cuda:
        return obj.device
    if not torch.cuda.is_available():
        raise RuntimeError("mrrt.nufft_amd.dtft runs on a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _as_dev(a, dev, dtype):
    if isinstance(a, torch.Tensor):
        return a.detach().to(dev).to(dtype)
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev).to(dtype)


def _grid_coords(shape, n_shift, dev):
    """(prod(shape), ndim) float64 coordinates n - n_shift, Fortran order."""
    axes = [torch.arange(int(s), dtype=torch.float64, device=dev) - float(n_shift[d]) for d, s in enumerate(shape)]
    mesh = torch.meshgrid(*axes, indexing="ij")
    return torch.stack([g.permute(*reversed(range(len(shape)))).reshape(-1) for g in mesh], dim=1)


def _f_flatten(t, nd):
    """Fortran ravel of the first nd axes: (prod(shape[:nd]), rest)."""
    lead = t.permute(*reversed(range(nd)), *range(nd, t.dim())) if nd > 1 else t
    return lead.reshape(-1, int(np.prod(t.shape[nd:])) if t.dim() > nd else 1)


def _shape_and_n_shift(a_ndim, a_shape, omega_cols, shape, n_shift):
    if shape is None:
        if a_ndim == omega_cols + 1:
            shape = tuple(a_shape[:-1])
        elif a_ndim == omega_cols:
            shape = tuple(a_shape)
        else:
            raise ValueError("shape must be specified")
    shape = tuple(int(s) for s in np.atleast_1d(shape))
    if len(shape) == omega_cols + 1:   # multiple images
        shape = shape[:-1]
    if len(shape) != omega_cols:
        raise ValueError("length of shape must match number of columns in omega")
    if n_shift is None:
        n_shift = (0.0,) * omega_cols
    elif np.isscalar(n_shift):
        n_shift = (float(n_shift),) * omega_cols
    n_shift = tuple(float(v) for v in np.atleast_1d(np.squeeze(np.asarray(n_shift, dtype=float))))
    if len(n_shift) != omega_cols:
        raise ValueError("must specify one shift per axis")
    return shape, n_shift


def _chunk(m, n, useloop):
    limit = (1 << 22) if useloop else (1 << 26)
    return max(1, min(m, limit // max(n, 1)))


def dtft(x, omega, shape=None, n_shift=None, useloop=False, xp=None, device=None):
    """Exact forward NUDFT of x (``shape`` or ``shape + (L,)``, Fortran
    order, or raveled with ``shape`` given) at the rows of omega (M, ndim);
    returns (M,) or (M, L).  ``xp`` is accepted for the reference's
    signature (numpy in, numpy out; tensors stay tensors)."""
    was_numpy = not isinstance(x, torch.Tensor)
    dev = _device(x if not was_numpy else omega, device)
    om = _as_dev(omega, dev, torch.float64)
    if om.dim() != 2:
        raise ValueError("omega must be 2d")
    xa = x if isinstance(x, torch.Tensor) else np.asarray(x)
    shape, n_shift = _shape_and_n_shift(xa.ndim, xa.shape, om.shape[1], shape, n_shift)
    npix = int(np.prod(shape))
    xt = _as_dev(xa, dev, torch.complex128)
    if xt.numel() % npix != 0:
        raise ValueError("bad input signal size")
    if xt.dim() >= len(shape) and tuple(xt.shape[:len(shape)]) == shape:
        xt = _f_flatten(xt, len(shape))
    else:
        xt = xt.reshape(-1, 1) if xt.dim() == 1 else xt.reshape(npix, -1)
    nn = _grid_coords(shape, n_shift, dev)
    out = torch.empty((om.shape[0], xt.shape[1]), dtype=torch.complex128, device=dev)
    c = _chunk(om.shape[0], npix, useloop)
    for s in range(0, om.shape[0], c):
        arg = om[s:s + c] @ nn.t()
        out[s:s + c] = torch.polar(torch.ones_like(arg), -arg) @ xt
    if out.shape[1] == 1:
        out = out[:, 0]
    return out.cpu().numpy() if was_numpy else out


def dtft_adj(xk, omega, shape=None, n_shift=None, useloop=False, xp=None, device=None):
    """Exact adjoint NUDFT of xk ((M,) or (M, L)) onto the grid ``shape``
    (Fortran order; the reference infers ``shape`` from xk's dimensions when
    it is not given); returns ``shape`` or ``shape + (L,)``."""
    was_numpy = not isinstance(xk, torch.Tensor)
    dev = _device(xk if not was_numpy else omega, device)
    om = _as_dev(omega, dev, torch.float64)
    xa = xk if isinstance(xk, torch.Tensor) else np.asarray(xk)
    shape, n_shift = _shape_and_n_shift(xa.ndim, xa.shape, om.shape[1], shape, n_shift)
    npix = int(np.prod(shape))
    X = _as_dev(xa, dev, torch.complex128).reshape(om.shape[0], -1)
    nn = _grid_coords(shape, n_shift, dev)
    out = torch.zeros((npix, X.shape[1]), dtype=torch.complex128, device=dev)
    c = _chunk(om.shape[0], npix, useloop)
    for s in range(0, om.shape[0], c):
        arg = nn @ om[s:s + c].t()
        out += torch.polar(torch.ones_like(arg), arg) @ X[s:s + c]
    L = out.shape[1]
    out = out.reshape(tuple(reversed(shape)) + (L,)).permute(*reversed(range(len(shape))), len(shape))
    if L == 1:
        out = out[..., 0]
    return out.cpu().numpy() if was_numpy else out
