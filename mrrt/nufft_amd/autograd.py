"""Differentiable table1 NUFFT (torch.autograd.Function).

``nufft(op, x, omega)`` computes y = A(omega) x with the operator ``op`` (a
``NufftBase``) and is differentiable with respect to the image ``x`` and the
sampling points ``omega``:

* d/dx is the adjoint transform (the same HIP scatter kernels as ``adj``);
* d/domega is a derivative-of-table gather (mi355nufft_interp_grad): the
  piecewise-linear table interpolant is differentiated exactly, and the
  derivative of phase_after (exp(i omega . (n_shift - n_mid))) is included.

``nufft_adjoint(op, y, omega)`` is x = A(omega)^H y, differentiable with
respect to ``y`` (forward transform) and ``omega`` (same gather kernel).

Layouts are torch-native (C order, repetitions leading): x is (*Nd) or
(C, *Nd); y is (M,) or (C, M); omega is (M, ndim) real on the device.
"""

import ctypes

import numpy as np
import torch

from . import _lib

__all__ = ["nufft", "nufft_adjoint"]


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _image_view(op, xb):
    """(C, *Nd) tensor -> logical (*Nd, C) view with its own strides."""
    return xb.permute(*(tuple(range(1, op.ndim + 1)) + (0,)))


def _points_key(omega):
    """Identity of a points tensor, including torch's in-place version counter."""
    if omega is None:
        return None
    return (omega.data_ptr(), omega._version, tuple(omega.shape), omega.dtype)


def _ensure_points(op, omega, key):
    if omega is None:
        return
    if getattr(op, "_ad_points_key", None) != key or op._points_version != getattr(op, "_ad_points_version", -1):
        op.set_points(omega.detach())
        op._ad_points_key = key
        op._ad_points_version = op._points_version


def _forward_grid(op, xb):
    """fftn(x * sn, Kd) of a (C, *Nd) batch; returns the (C, prod Kd) grid."""
    C = xb.shape[0]
    grid = torch.empty((C, int(np.prod(op.Kd))), dtype=op._tcdt, device=op.device)
    xv = _image_view(op, xb)
    op._sync_stream()
    _lib.check(op._lib.mi355nufft_forward_grid(op._plan, _vp(xv), _lib.i64arr(list(xv.stride())), _vp(grid), C))
    return grid


def _interp(op, grid, C):
    y = torch.empty((C, op.M), dtype=op._tcdt, device=op.device)
    _lib.check(op._lib.mi355nufft_interp(op._plan, _vp(grid), _vp(y), C))
    return y


def _adjoint_image(op, y):
    """A^H y for y (C, M) contiguous, without adjoint_scalefactor; (C, *Nd)."""
    C = y.shape[0]
    out = torch.empty((C,) + tuple(op.Nd), dtype=op._tcdt, device=op.device)
    op._adjoint_dev(y, C, out=_image_view(op, out))
    if op.adjoint_scalefactor != 1:
        out = out / op.adjoint_scalefactor
    return out


def _omega_grad(op, grid, w, omega):
    C = w.shape[0]
    g = torch.empty((op.M, op.ndim), dtype=op._trdt, device=op.device)
    _lib.check(op._lib.mi355nufft_interp_grad(op._plan, _vp(grid), _vp(w), _vp(g), C))
    return g.to(omega.dtype)


class _Forward(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, omega, op):
        key = _points_key(omega)
        _ensure_points(op, omega, key)
        batched = x.dim() == op.ndim + 1
        xb = (x if batched else x.unsqueeze(0)).to(op._tcdt)
        grid = _forward_grid(op, xb)
        y = _interp(op, grid, xb.shape[0])
        ctx.op, ctx.batched, ctx.key = op, batched, key
        ctx.save_for_backward(grid, omega if omega is not None else torch.empty(0))
        ctx.has_omega = omega is not None
        return y if batched else y[0]

    @staticmethod
    def backward(ctx, gy):
        op = ctx.op
        grid, omega = ctx.saved_tensors
        if ctx.has_omega:
            _ensure_points(op, omega, ctx.key)
        w = (gy if ctx.batched else gy.unsqueeze(0)).to(op._tcdt).contiguous()
        gx = gw = None
        if ctx.needs_input_grad[0]:
            gx = _adjoint_image(op, w)
            if not ctx.batched:
                gx = gx[0]
        if ctx.has_omega and ctx.needs_input_grad[1]:
            gw = _omega_grad(op, grid, w, omega)
        return gx, gw, None


class _Adjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, omega, op):
        key = _points_key(omega)
        _ensure_points(op, omega, key)
        batched = y.dim() == 2
        yb = (y if batched else y.unsqueeze(0)).to(op._tcdt).contiguous()
        x = _adjoint_image(op, yb)
        ctx.op, ctx.batched, ctx.key = op, batched, key
        ctx.save_for_backward(yb, omega if omega is not None else torch.empty(0))
        ctx.has_omega = omega is not None
        return x if batched else x[0]

    @staticmethod
    def backward(ctx, gx):
        op = ctx.op
        yb, omega = ctx.saved_tensors
        if ctx.has_omega:
            _ensure_points(op, omega, ctx.key)
        gxb = (gx if ctx.batched else gx.unsqueeze(0)).to(op._tcdt)
        grid = _forward_grid(op, gxb)
        gy = gw = None
        if ctx.needs_input_grad[0]:
            gy = _interp(op, grid, gxb.shape[0])
            if not ctx.batched:
                gy = gy[0]
        if ctx.has_omega and ctx.needs_input_grad[1]:
            # d/domega Re<gx, A^H y> = d/domega Re<A gx, y>: the gather gradient with w = y
            gw = _omega_grad(op, grid, yb, omega)
        return gy, gw, None


def nufft(op, x, omega=None):
    """Differentiable forward NUFFT y = A(omega) x (see module docstring)."""
    return _Forward.apply(x, omega, op)


def nufft_adjoint(op, y, omega=None):
    """Differentiable adjoint NUFFT x = A(omega)^H y (see module docstring)."""
    return _Adjoint.apply(y, omega, op)
