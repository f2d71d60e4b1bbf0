"""Pruned, deapodization-fused FFT passes (power-of-two grids) against numpy's
FFT in float64 and against the rocFFT path of the same plan.

forward_grid(x) = fftn(x * sn, Kd)                       (_nufft.py:1326-1331)
adjoint_image(G) = ifftn(G)[:Nd] * prod(Kd) * asf * sn    (_nufft.py:1524-1572)
"""

import ctypes

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase, _lib
from mrrt.nufft_amd import trajectories as trj
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

# float32 FFT of length K: rounding grows like log2(K) * eps
TOL = {"single": 2e-6, "double": 1e-13}

GEOMS = [
    ((12,), (16,)),
    ((700,), (2048,)),
    ((16, 8), (16, 32)),
    ((256, 256), (512, 512)),
    ((31, 17), (64, 2048)),
    ((64, 64), (64, 64)),          # no padding
    ((10, 12, 20), (16, 32, 64)),
    ((32, 32, 32), (32, 32, 32)),
]


def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


def _op(Nd, Kd, precision, **kw):
    M = 64
    om = np.random.default_rng(0).uniform(-np.pi, np.pi, (M, len(Nd)))
    return NufftBase(Nd, om, Jd=[min(6, k) for k in Kd], Kd=Kd, Ld=64, precision=precision, device=dev(), **kw)


def _image(Nd, nreps, cdt, seed):
    """(nreps, N_{d-1}, ..., N_0) C-contiguous = Fortran Nd per repetition."""
    xn = trj.complex_gaussian((nreps,) + tuple(reversed(Nd)), seed=seed)
    return xn, torch.from_numpy(xn).to(dev()).to(cdt)


def _strides(t):
    # element strides of axes 0..ndim-1 (reversed tensor dims) and the repetition
    s = list(t.stride())
    return [s[i] for i in range(len(s) - 1, 0, -1)] + [s[0]]


def _forward_grid(A, xt, nreps):
    g = torch.empty((nreps, int(np.prod(A.Kd))), dtype=A._tcdt, device=dev())
    A._sync_stream()
    _lib.check(A._lib.mi355nufft_forward_grid(A._plan, ctypes.c_void_p(xt.data_ptr()),
                                              _lib.i64arr(_strides(xt)), ctypes.c_void_p(g.data_ptr()), nreps))
    return g


def _adjoint_image(A, g, nreps):
    x = torch.empty((nreps,) + tuple(reversed(A.Nd)), dtype=A._tcdt, device=dev())
    A._sync_stream()
    _lib.check(A._lib.mi355nufft_adjoint_image(A._plan, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                                               _lib.i64arr(_strides(x)), nreps))
    return x


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("Nd,Kd", GEOMS)
def test_forward_grid_vs_numpy(precision, Nd, Kd):
    A = _op(Nd, Kd, precision)
    assert A.plan_info()["fused_fft"] == (precision == "single" or max(Kd) <= 1024)
    nreps = 3
    xn, xt = _image(Nd, nreps, A._tcdt, seed=5)
    g = _forward_grid(A, xt, nreps).cpu().numpy()
    sn = A.sn.cpu().numpy().astype(np.float64)
    for r in range(nreps):
        ref = np.fft.fftn(xn[r].T * sn, s=Kd)                    # natural axis order
        got = g[r].reshape(tuple(reversed(Kd))).T
        assert rel_l2(got, ref) <= TOL[precision], (r, rel_l2(got, ref))


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("Nd,Kd", GEOMS)
def test_adjoint_image_vs_numpy(precision, Nd, Kd):
    A = _op(Nd, Kd, precision, adjoint_scalefactor=0.75)
    nreps = 2
    gn = trj.complex_gaussian((nreps, int(np.prod(Kd))), seed=6)
    g = torch.from_numpy(gn).to(dev()).to(A._tcdt)
    x = _adjoint_image(A, g, nreps).cpu().numpy()
    sn = A.sn.cpu().numpy().astype(np.float64)
    sl = tuple(slice(0, n) for n in Nd)
    for r in range(nreps):
        ref = np.fft.ifftn(gn[r].reshape(tuple(reversed(Kd))).T)[sl] * np.prod(Kd) * 0.75 * sn
        assert rel_l2(x[r].T, ref) <= TOL[precision], (r, rel_l2(x[r].T, ref))


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("Nd,Kd,ortho", [((64, 48), (128, 128), False), ((20, 24, 16), (32, 64, 32), True),
                                         ((100,), (256,), False)])
def test_operator_fused_vs_rocfft(precision, Nd, Kd, ortho):
    """Whole transforms with the fused FFT equal the rocFFT path of the same plan."""
    om = trj.golden_angle_2d(24, 64) if len(Nd) == 2 else np.random.default_rng(1).uniform(
        -np.pi, np.pi, (500, len(Nd)))
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=256, precision=precision, ortho=ortho, device=dev())
    nreps = 2
    x = torch.from_numpy(trj.complex_gaussian(Nd + (nreps,), seed=7)).to(dev()).to(A._tcdt)
    y1 = A._forward_dev(x, nreps).clone()
    xa1 = A._adjoint_dev(y1, nreps).clone()
    assert A.plan_info()["fused_fft"]
    A.set_fft("rocfft")
    assert not A.plan_info()["fused_fft"]
    y2 = A._forward_dev(x, nreps).clone()
    xa2 = A._adjoint_dev(y1, nreps).clone()
    A.set_fft("auto")
    tol = 3 * TOL[precision]
    assert rel_l2(y1.cpu().numpy(), y2.cpu().numpy()) <= tol
    assert rel_l2(xa1.cpu().numpy(), xa2.cpu().numpy()) <= tol


def test_non_power_of_two_grid_uses_rocfft():
    """K = 48 has no fused passes; the transforms still run (rocFFT) and match
    numpy."""
    A = _op((20, 30), (48, 64), "double")
    assert not A.plan_info()["fused_fft"]
    xn, xt = _image((20, 30), 1, A._tcdt, seed=8)
    g = _forward_grid(A, xt, 1).cpu().numpy()
    ref = np.fft.fftn(xn[0].T * A.sn.cpu().numpy(), s=(48, 64))
    assert rel_l2(g[0].reshape(64, 48).T, ref) <= 1e-13


def test_set_fft_rejects_bad_method():
    A = _op((16,), (32,), "single")
    with pytest.raises(KeyError):
        A.set_fft("cufft")
    assert A._lib.mi355nufft_set_fft(A._plan, 7) == 1
