"""Device NUDFT (dtft / dtft_adj) and the precision setter of the drop-in
(mrrt/nufft/_dtft.py:16-216, mrrt/nufft/_nufft.py:520-545), against the
oracle."""

import numpy as np
import pytest
import torch

import mrrt.nufft_amd as M
from mrrt.nufft_amd import trajectories as trj
from oracle import nufft_oracle as O
from tests.cases import rel_l2, uniform_freqs

pytestmark = pytest.mark.gpu


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


@pytest.mark.parametrize("Nd,n_shift", [((16,), (5,)), ((8, 12), (1, 3)), ((8, 12, 10), (1, 3, 2))])
def test_dtft_vs_oracle_and_fft(Nd, n_shift):
    rs = np.random.RandomState(1234)
    x = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    om = trj.kooshball_3d(8, 10, seed=1)[:, :len(Nd)] * 0.9 if len(Nd) > 1 else np.linspace(-3, 3, 50)[:, None]
    X = M.dtft(x, om, n_shift=n_shift)
    assert isinstance(X, np.ndarray) and X.shape == (om.shape[0],)
    Xo = O.dtft(x.ravel(order="F"), om, Nd, np.asarray(n_shift, float))
    assert rel_l2(X, Xo) <= 1e-12
    # uniform frequencies: the DTFT is the shifted FFT (test_dtft.py:39-115)
    omu = uniform_freqs(Nd)
    Xu = M.dtft(x, omu, n_shift=n_shift)
    Xf = np.fft.fftn(x).ravel(order="F") * np.exp(1j * (omu @ np.asarray(n_shift, float)))
    np.testing.assert_allclose(Xu, Xf, atol=1e-10)
    # adjoint: the raveled pair (M,) -> shape, and with repetitions
    y = rs.standard_normal(om.shape[0]) + 1j * rs.standard_normal(om.shape[0])
    xa = M.dtft_adj(y, om, shape=Nd, n_shift=n_shift)
    assert xa.shape == Nd
    xo = O.dtft_adj(y, om, Nd, np.asarray(n_shift, float)).reshape(Nd, order="F")
    assert rel_l2(xa, xo) <= 1e-12
    Y = np.stack([y, 2 * y], axis=-1)
    xa2 = M.dtft_adj(Y, om, shape=Nd, n_shift=n_shift)
    assert xa2.shape == Nd + (2,)
    assert rel_l2(xa2[..., 1], 2 * xo) <= 1e-12
    # device tensors stay on the device
    Xt = M.dtft(torch.from_numpy(x).to(dev()), torch.from_numpy(om).to(dev()), n_shift=n_shift)
    assert isinstance(Xt, torch.Tensor) and Xt.is_cuda
    assert rel_l2(Xt.cpu().numpy(), Xo) <= 1e-12


def test_dtft_repetitions_and_useloop():
    Nd = (6, 5)
    rs = np.random.RandomState(7)
    x = rs.standard_normal(Nd + (3,)) + 1j * rs.standard_normal(Nd + (3,))
    om = trj.golden_angle_2d(5, 8)
    X = M.dtft(x, om)
    assert X.shape == (om.shape[0], 3)
    for r in range(3):
        assert rel_l2(X[:, r], O.dtft(x[..., r].ravel(order="F"), om, Nd)) <= 1e-12
    np.testing.assert_allclose(M.dtft(x, om, useloop=True), X, rtol=0, atol=1e-12)


@pytest.mark.parametrize("mode", ["table", "sparse"])
def test_precision_setter_rebuilds_operator(mode):
    Nd, Kd = (20, 24), (40, 48)
    om = trj.golden_angle_2d(10, 40)
    A = M.NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, mode=mode, device=dev())
    x = trj.complex_gaussian(Nd, seed=3)
    for precision, tol in (("double", 1e-12), ("single", 1e-5), ("double", 1e-12)):
        A.precision = precision
        assert A.precision == precision
        B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=512, precision=precision, mode=mode)
        y = A.fft(x)
        assert y.dtype == B.cdt
        assert rel_l2(y, B.fft(x)) <= tol
        assert rel_l2(A.adj(y), B.adj(y)) <= tol
    with pytest.raises(ValueError):
        A.precision = "half"
    assert A.precision == "double"


def test_numpy_results_are_ordinary_arrays():
    """numpy in, numpy out: results are new arrays that own their memory
    (not views of page-locked staging), in the reference's Fortran layout,
    equal to the device-tensor path; repeated calls do not alias."""
    Nd, Kd = (40, 36), (80, 72)
    om = trj.golden_angle_2d(30, 80)
    A = M.NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    x = np.asfortranarray(trj.complex_gaussian(Nd + (3,), seed=8).astype(np.complex64))
    y1 = A.fft(x)
    y2 = A.fft(2 * x)
    assert isinstance(y1, np.ndarray) and y1.flags.owndata and y1.dtype == np.complex64
    assert y1.shape == (om.shape[0], 3) and y1.flags.f_contiguous
    np.testing.assert_allclose(y2, 2 * y1, rtol=1e-5, atol=1e-6 * np.abs(y1).max())
    yd = A.fft(torch.from_numpy(np.ascontiguousarray(x)).to(dev()))
    assert rel_l2(y1, yd.cpu().numpy()) <= 1e-6
    xa = A.adj(y1)
    assert xa.flags.owndata and xa.flags.f_contiguous and xa.shape == Nd + (3,)
    xa_c = A.adj(np.ascontiguousarray(y1))
    assert rel_l2(xa, xa_c) <= 1e-6


def test_measure_atomics_kinds():
    """mi355nufft_measure_atomics: every kind (LDS f64 / u32 / u64 at
    scattered addresses, coalesced global f32) returns a finite positive
    rate; other kinds are rejected on the host before any launch."""
    import ctypes
    from mrrt.nufft_amd import _lib
    dev()
    lib = _lib.get_lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rates = {}
    for kind in (0, 1, 2, 3):
        v = ctypes.c_double(0.0)
        _lib.check(lib.mi355nufft_measure_atomics(kind, ctypes.byref(v), stream))
        assert np.isfinite(v.value) and v.value > 0, kind
        rates[kind] = v.value
    # the device's LDS does many more lane operations per second than global atomics
    assert min(rates[0], rates[1], rates[3]) > rates[2]
    for kind in (-1, 4):
        v = ctypes.c_double(0.0)
        assert lib.mi355nufft_measure_atomics(kind, ctypes.byref(v), stream) != 0
