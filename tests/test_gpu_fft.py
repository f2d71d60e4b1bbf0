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


def _adjoint_call(A, y, nreps, grid):
    out = torch.empty((nreps,) + tuple(reversed(A.Nd)), dtype=A._tcdt, device=dev())
    A._sync_stream()
    gp = None if grid is None else ctypes.c_void_p(grid.data_ptr())
    _lib.check(A._lib.mi355nufft_adjoint(A._plan, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                         _lib.i64arr(_strides(out)), nreps, gp))
    return out


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("Nd,Kd", [((48, 40), (64, 128)), ((12, 10, 14), (16, 32, 32)), ((90,), (256,))])
def test_plan_grid_kept_zero_between_adjoints(precision, Nd, Kd):
    """Adjoints through the plan-owned grid (cleared by the previous adjoint's
    fused passes) equal adjoints through a fresh caller grid (cleared by the
    scatter): repeated calls, a smaller then a larger batch, forwards in
    between, and a switch to rocFFT and back."""
    om = np.random.default_rng(3).uniform(-np.pi, np.pi, (700, len(Nd)))
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=256, precision=precision, device=dev())
    nk = int(np.prod(Kd))
    tol = 3 * TOL[precision]
    for step, (nreps, fft) in enumerate([(2, "auto"), (2, "auto"), (1, "auto"), (3, "auto"), (2, "rocfft"),
                                         (2, "auto"), (2, "auto")]):
        A.set_fft(fft)
        y = torch.from_numpy(trj.complex_gaussian((nreps, A.M), seed=20 + step)).to(dev()).to(A._tcdt)
        ref = _adjoint_call(A, y, nreps, torch.full((nreps, nk), 7.0, dtype=A._tcdt, device=dev()))
        got = _adjoint_call(A, y, nreps, None)
        assert rel_l2(got.cpu().numpy(), ref.cpu().numpy()) <= tol, (step, nreps, fft)
        if step % 2:
            xn, xt = _image(Nd, nreps, A._tcdt, seed=30 + step)
            strides = _strides(xt)
            yy = torch.empty((nreps, A.M), dtype=A._tcdt, device=dev())
            _lib.check(A._lib.mi355nufft_forward(A._plan, ctypes.c_void_p(xt.data_ptr()), _lib.i64arr(strides),
                                                 ctypes.c_void_p(yy.data_ptr()), nreps, None))


def test_plan_grid_adjoint_under_graph_capture():
    """A captured adjoint through the plan-owned grid clears it itself; replays
    with eager adjoints in between all equal the eager result."""
    Nd, Kd = (64, 64), (128, 128)
    A = NufftBase(Nd, trj.golden_angle_2d(40, 128), Jd=6, Kd=Kd, Ld=1024, device=dev())
    y = torch.from_numpy(trj.complex_gaussian((1, A.M), seed=40)).to(dev()).to(A._tcdt)
    ref = A._adjoint_dev(y, 1).clone()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        A._adjoint_dev(y, 1)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = A._adjoint_dev(y, 1)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        assert rel_l2(out.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
        e = A._adjoint_dev(y, 1)
        assert rel_l2(e.cpu().numpy(), ref.cpu().numpy()) <= 1e-6
