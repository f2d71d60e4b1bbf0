"""GPU parity: the HIP path against the oracle, the reference's golden vectors
and the exact NUDFT.

Bars (written in each test): relative L2 error <= 1e-5 for complex64 and
<= 1e-12 for complex128 against the reference restatement; the reference's
own tolerances against the exact DTFT (mrrt/nufft/tests/test_nufft.py).
"""

import ctypes

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase, _lib
from mrrt.nufft_amd import trajectories as trj
from oracle import nufft_oracle as O
from tests.cases import golden_cases, perturbed_gridpoints, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# ------------------------------------------------------------- raw kernels
@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["key"])
def test_table1_kernels_vs_reference_golden(case):
    """mi355nufft_table1_forward/adjoint == the reference C kernels' outputs."""
    lib = _lib.get_lib()
    ndim, Kd, Jd, L = case["ndim"], case["Kd"], case["Jd"], case["L"]
    single = case["tm"].dtype == np.float32
    prec = 0 if single else 1
    hs = [_t(h) for h in case["hs"]]
    hp = (ctypes.c_void_p * ndim)(*[h.data_ptr() for h in hs])
    tm = _t(case["tm"].T)                       # (ndim, M)
    ck = _t(case["ck"].T)                       # (nreps, nk)
    M = case["tm"].shape[0]
    nreps = case["ck"].shape[1]
    fm = torch.empty((nreps, M), dtype=ck.dtype, device=dev())
    _lib.check(lib.mi355nufft_table1_forward(ndim, _lib.i64arr(Kd), _lib.i32arr(Jd), L, prec,
                                             int(case["complex_kernel"]), ctypes.c_void_p(ck.data_ptr()), hp,
                                             ctypes.c_void_p(tm.data_ptr()), M, ctypes.c_void_p(fm.data_ptr()),
                                             nreps, _stream()))
    fin = _t(case["fm"].T)
    out = torch.empty((nreps, int(np.prod(Kd))), dtype=ck.dtype, device=dev())
    _lib.check(lib.mi355nufft_table1_adjoint(ndim, _lib.i64arr(Kd), _lib.i32arr(Jd), L, prec,
                                             int(case["complex_kernel"]), ctypes.c_void_p(fin.data_ptr()), hp,
                                             ctypes.c_void_p(tm.data_ptr()), M, ctypes.c_void_p(out.data_ptr()),
                                             nreps, _stream()))
    torch.cuda.synchronize()
    fwd = fm.cpu().numpy().T
    adj = out.cpu().numpy().T
    tol = 2e-6 if single else 1e-14
    assert rel_l2(fwd, case["fwd"]) <= tol
    assert rel_l2(adj, case["adj"]) <= tol
    if not case["complex_kernel"]:
        # same operations, same order, no contraction: bit-identical
        np.testing.assert_array_equal(fwd, case["fwd"])


# ------------------------------------------------------- operator vs oracle
def _pair(Nd, om, **kw):
    A = NufftBase(Nd, om, device=dev(), **kw)
    B = O.NufftOracle(Nd, om, **kw)
    return A, B


CASES_2D = [
    # (Nd, Kd, Jd, Ld, n_shift)
    ((64, 64), (128, 128), 6, 1024, None),
    ((32, 24), (64, 49), (6, 7), 512, (3.3, -2.0)),
    ((33, 31), (66, 62), 5, 1024, (16, 15)),
]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("Nd,Kd,Jd,Ld,n_shift", CASES_2D)
def test_operator_2d_vs_oracle(precision, phasing, Nd, Kd, Jd, Ld, n_shift):
    om = trj.golden_angle_2d(40, 2 * Nd[0])
    A, B = _pair(Nd, om, Jd=Jd, Kd=Kd, Ld=Ld, n_shift=n_shift, precision=precision, phasing=phasing)
    x = trj.complex_gaussian(Nd, seed=1)
    y = A.fft(x)
    yo = B.fft(x)
    assert y.dtype == B.cdt
    assert rel_l2(y, yo) <= TOL[precision]
    xa = A.adj(yo)
    assert rel_l2(xa, B.adj(yo)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_operator_1d_3d_vs_oracle(precision, phasing):
    om = np.random.default_rng(2).uniform(-np.pi, np.pi, (500, 1))
    A, B = _pair((64,), om, Jd=6, Kd=(128,), Ld=1024, n_shift=(5.5,), precision=precision, phasing=phasing)
    x = trj.complex_gaussian((64,), seed=3)
    assert rel_l2(A.fft(x), B.fft(x)) <= TOL[precision]
    y = trj.complex_gaussian((500,), seed=4)
    assert rel_l2(A.adj(y), B.adj(y)) <= TOL[precision]
    om = trj.kooshball_3d(64, 32, seed=1)
    Nd, Kd = (16, 12, 10), (32, 24, 20)
    A, B = _pair(Nd, om, Jd=(6, 5, 4), Kd=Kd, Ld=512, precision=precision, phasing=phasing)
    x = trj.complex_gaussian(Nd, seed=5)
    assert rel_l2(A.fft(x), B.fft(x)) <= TOL[precision]
    y = trj.complex_gaussian((om.shape[0],), seed=6)
    assert rel_l2(A.adj(y), B.adj(y)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("Nd,Kd,nreps", [((24, 20), (48, 40), 13), ((12, 10, 8), (24, 20, 16), 9), ((40,), (80,), 17)])
def test_many_repetitions_vs_oracle(precision, phasing, Nd, Kd, nreps):
    """Repetitions beyond the gather's staging chunk (8 per pass): chunk
    boundaries, a partial last chunk, 1-D / 2-D / 3-D."""
    om = np.random.default_rng(11).uniform(-np.pi, np.pi, (400, len(Nd)))
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=512, precision=precision, phasing=phasing)
    x = trj.complex_gaussian(Nd + (nreps,), seed=12)
    y = A.fft(x)
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    yb = trj.complex_gaussian((om.shape[0], nreps), seed=13)
    assert rel_l2(A.adj(yb), B.adj(yb)) <= TOL[precision]
    for r in (0, 7, 8, nreps - 1):  # each repetition on its own
        assert rel_l2(y[:, r], B.fft(x[..., r])) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("ortho", [False, True])
@pytest.mark.parametrize("Nd,Kd", [((48, 40), (64, 128)), ((12, 10, 8), (32, 16, 16))])
def test_fft_pass_phase_vs_oracle(precision, phasing, ortho, Nd, Kd):
    """Power-of-two grids: the forward's last FFT pass applies phase_before
    and the ortho scale, the adjoint's first pass conj(phase_before).  Both
    against the oracle and against the rocFFT path of the same plan (where the
    gather and the scatter apply them)."""
    om = np.random.default_rng(21).uniform(-np.pi, np.pi, (900, len(Nd)))
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=512, n_shift=tuple(1.5 for _ in Nd), ortho=ortho, precision=precision,
                 phasing=phasing)
    assert A.plan_info()["fused_fft"]
    x = trj.complex_gaussian(Nd + (2,), seed=22)
    y = A.fft(x)
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    yb = trj.complex_gaussian((om.shape[0], 2), seed=23)
    xa = A.adj(yb)
    assert rel_l2(xa, B.adj(yb)) <= TOL[precision]
    A.set_fft("rocfft")
    assert rel_l2(A.fft(x), y) <= TOL[precision]
    assert rel_l2(A.adj(yb), xa) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_operator_options_vs_oracle(precision):
    """ortho, adjoint_scalefactor, repetitions, order='C', real input."""
    Nd, Kd = (32, 32), (64, 64)
    om = trj.radial_2d(32, 64)
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, ortho=True, adjoint_scalefactor=0.7, precision=precision)
    x = trj.complex_gaussian(Nd + (3,), seed=7)
    y = A.fft(x)
    assert y.shape == (om.shape[0], 3)
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    xa = A.adj(y)
    assert xa.shape == Nd + (3,)
    assert rel_l2(xa, B.adj(y)) <= TOL[precision]
    # order='C': repetitions leading (_nufft.py:415-423, :443-451)
    C = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, order="C", device=dev())
    B2 = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
    xc = np.moveaxis(x, -1, 0).copy()
    yc = C.fft(xc)
    assert yc.shape == (3, om.shape[0])
    assert rel_l2(yc.T, B2.fft(x)) <= TOL[precision]
    xac = C.adj(yc)
    assert xac.shape == (3,) + Nd
    assert rel_l2(np.moveaxis(xac, 0, -1), B2.adj(yc.T)) <= TOL[precision]
    # real-valued input is promoted to complex (_nufft.py:1312)
    xr = np.random.default_rng(0).standard_normal(Nd)
    assert rel_l2(C.fft(xr), B2.fft(xr)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_many_lines_c_order_vs_oracle(precision):
    """Repetitions leading over C-ordered images with K0 = 512 and 2048 rows:
    the axis-0 FFT passes take up to 16 neighbouring rows per workgroup."""
    Nd, Kd, R = (256, 16), (512, 32), 128
    om = np.random.default_rng(21).uniform(-np.pi, np.pi, (1500, 2))
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, order="C", device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
    x = trj.complex_gaussian((R,) + Nd, seed=22)
    y = A.fft(x)
    assert y.shape == (R, om.shape[0])
    assert rel_l2(y.T, B.fft(np.moveaxis(x, 0, -1))) <= TOL[precision]
    xa = A.adj(y)
    assert rel_l2(np.moveaxis(xa, 0, -1), B.adj(y.T.astype(B.cdt))) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_c1_radial_vs_oracle_and_nudft(precision):
    """Nd 128^2, Kd 256^2, J 6, L 1024, 64 radial spokes x 256 readout."""
    Nd, Kd = (128, 128), (256, 256)
    om = trj.radial_2d(64, 256)
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
    x = trj.complex_gaussian(Nd, seed=0)
    y = A.fft(x)
    yo = B.fft(x)
    assert rel_l2(y, yo) <= TOL[precision]
    xa = A.adj(yo)
    assert rel_l2(xa, B.adj(yo)) <= TOL[precision]
    # exact NUDFT on a seeded subset of samples / pixels
    sel = np.random.default_rng(1).choice(om.shape[0], 512, replace=False)
    ye = O.dtft(x.ravel(order="F"), om[sel], Nd, np.zeros(2))
    assert rel_l2(y[sel], ye) < 2e-3


def test_torch_tensor_io_and_layouts():
    """Device tensors stay on the device; strided inputs are read in place."""
    Nd, Kd = (32, 40), (64, 80)
    om = trj.golden_angle_2d(16, 64)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024)
    x = trj.complex_gaussian(Nd, seed=9).astype(np.complex64)
    xt = torch.from_numpy(x).to(dev())
    y = A.fft(xt)
    assert isinstance(y, torch.Tensor) and y.is_cuda
    assert rel_l2(y.cpu().numpy(), B.fft(x)) <= 1e-5
    xtt = torch.from_numpy(np.asfortranarray(x)).to(dev())      # different strides, same values
    assert rel_l2(A.fft(xtt).cpu().numpy(), B.fft(x)) <= 1e-5
    xa = A.adj(y)
    assert xa.shape == Nd and xa.is_cuda
    assert rel_l2(xa.cpu().numpy(), B.adj(y.cpu().numpy())) <= 1e-5


# ---------------------------------------------- reference test cases (DTFT)
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_nufft_adj_kat(phasing):
    """mrrt/nufft/tests/test_nufft.py:60-96."""
    o1 = 2 * np.pi * np.array([0.0, 0.1, 0.3, 0.4, 0.7, 0.9])
    om = np.stack((o1, o1[::-1].copy()), axis=-1)
    st = NufftBase(omega=om, Nd=(4, 8), Jd=[8, 8], Kd=(8, 16), Ld=1024, n_shift=[2.7, 3.1], phasing=phasing,
                   mode="table", device=dev())
    data = np.arange(1, 7) ** 2
    xd = O.dtft_adj(data, om, (4, 8), [2.7, 3.1]).reshape((4, 8), order="F")
    xn = st.adj(np.stack((data,) * 3, axis=-1))
    np.testing.assert_array_almost_equal(xd, xn[:, :, -1], decimal=4)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("order", ["F", "C"])
def test_nufft_1d_vs_dtft(precision, phasing, order):
    """mrrt/nufft/tests/test_nufft.py:99-169 (table mode)."""
    Nd, n_shift = 64, 32
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=6, Kd=128, n_shift=n_shift, mode="table", Ld=1024, precision=precision,
                  phasing=phasing, order=order, device=dev())
    x = rs.randn(Nd) + 1j * rs.randn(Nd)
    y = A.fft(x)
    y_true = O.dtft(x, om, (Nd,), [n_shift])
    np.testing.assert_allclose(y, y_true, rtol=1e-3, atol=1e-5)
    if order == "C":
        x_reps, sl1 = np.stack((x,) * 2, axis=0), (0, Ellipsis)
    else:
        x_reps, sl1 = np.stack((x,) * 2, axis=-1), (Ellipsis, 0)
    np.testing.assert_allclose(A.fft(x_reps)[sl1], y_true, rtol=1e-3, atol=1e-5)
    x_adj_true = O.dtft_adj(y, om, (Nd,), [n_shift])
    np.testing.assert_allclose(A.adj(y), x_adj_true, rtol=1e-3, atol=1e-5)
    y_reps = np.stack((y,) * 2, axis=0) if order == "C" else np.stack((y,) * 2, axis=-1)
    np.testing.assert_allclose(A.adj(y_reps)[sl1], x_adj_true, rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("Kd,Jd", [((32, 32), 6), ((33, 31), 7)])
@pytest.mark.parametrize("order", ["F", "C"])
def test_nufft_2d_vs_dtft(precision, phasing, Kd, Jd, order):
    """mrrt/nufft/tests/test_nufft.py:172-244 (table mode)."""
    Nd = (16, 16)
    n_shift = np.asarray(Nd) / 2
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=Jd, Kd=Kd, n_shift=n_shift, mode="table", Ld=1024, precision=precision,
                  phasing=phasing, order=order, device=dev())
    x = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    y = A.fft(x)
    y_true = O.dtft(x.ravel(order="F"), om, Nd, n_shift)
    np.testing.assert_allclose(y, y_true, rtol=1e-3, atol=1e-5)
    x_adj = A.adj(y)
    np.testing.assert_allclose(x_adj, O.dtft_adj(y, om, Nd, n_shift).reshape(Nd, order="F"), rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("order", ["F", "C"])
def test_nufft_3d_vs_dtft(precision, phasing, order):
    """mrrt/nufft/tests/test_nufft.py:247-324 (table mode)."""
    Nd, Kd = [8] * 3, [16] * 3
    n_shift = np.asarray(Nd) / 2
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=[6] * 3, Kd=Kd, n_shift=n_shift, mode="table", Ld=1024, precision=precision,
                  phasing=phasing, order=order, device=dev())
    x = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    y = A.fft(x)
    np.testing.assert_allclose(y, O.dtft(x.ravel(order="F"), om, tuple(Nd), n_shift), rtol=1e-2, atol=1e-4)
    if order == "C":
        x_reps, sl1 = np.stack((x,) * 4, axis=0), (0, Ellipsis)
    else:
        x_reps, sl1 = np.stack((x,) * 4, axis=-1), (Ellipsis, 0)
    np.testing.assert_allclose(A.fft(x_reps)[sl1], y, rtol=1e-5, atol=1e-7)
    x_adj = A.adj(y)
    np.testing.assert_allclose(x_adj, O.dtft_adj(y, om, tuple(Nd), n_shift).reshape(Nd, order="F"),
                               rtol=1e-2, atol=1e-4)


@pytest.mark.parametrize("mode", ["table", "sparse"])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_nufft_dtypes(precision, mode):
    """mrrt/nufft/tests/test_nufft.py:327-388 (in both modes)."""
    Nd, Kd, Jd, Ld = 64, 128, 6, 4096
    om = perturbed_gridpoints(Nd)
    A = NufftBase(omega=om, Nd=Nd, Jd=Jd, Kd=Kd, n_shift=Nd // 2, mode=mode, Ld=Ld, precision=precision,
                  device=dev())
    cdt = np.complex64 if precision == "single" else np.complex128
    A2 = NufftBase(omega=om.astype(A._real_dtype), Nd=Nd, Jd=Jd, Kd=Kd, n_shift=Nd // 2, mode=mode, Ld=Ld,
                   precision="auto", device=dev())
    assert A2._cplx_dtype == cdt
    x = np.random.randn(Nd) + 1j * np.random.randn(Nd)
    for xin in (x.astype(np.complex64), x.astype(np.complex128), x.real.astype(np.float32), x.real):
        assert A.fft(xin).dtype == cdt


# ------------------------------------------------------------ properties
@pytest.mark.parametrize("mode", ["table", "sparse"])
@pytest.mark.parametrize("Kd", [(96, 80), (64, 64)])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_adjointness(precision, Kd, mode):
    """<A x, y> = <x, A^H y> to rounding; (64, 64) runs the fused FFT passes
    with phase_before in them, (96, 80) rocFFT."""
    Nd = (48, 40)
    om = trj.golden_angle_2d(30, 96)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, mode=mode, device=dev())
    x = trj.complex_gaussian(Nd, seed=11)
    y = trj.complex_gaussian((om.shape[0],), seed=12)
    lhs = np.vdot(y, A.fft(x))
    rhs = np.vdot(A.adj(y), x) / np.prod(Kd) * np.prod(Kd)
    assert abs(lhs - rhs) <= (1e-5 if precision == "single" else 1e-12) * abs(lhs)


def test_edge_points_and_wrap():
    """Points on exact grid positions (the p == J*L/2 rounding edge), outside
    [-pi, pi) (periodic wrap) and M == 1."""
    Nd, Kd = (16, 16), (32, 32)
    g = 2 * np.pi / 32
    om = np.array([[0.0, 0.0], [g * 3, -g * 5], [-np.pi, -np.pi], [np.pi, 3 * np.pi], [-5 * np.pi + 0.1, 7.3],
                   [np.nextafter(np.float32(g), 0) * 1.0, 0.0]])
    for precision in ("single", "double"):
        A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
        x = trj.complex_gaussian(Nd, seed=13)
        assert rel_l2(A.fft(x), B.fft(x)) <= TOL[precision]
        y = trj.complex_gaussian((om.shape[0],), seed=14)
        assert rel_l2(A.adj(y), B.adj(y)) <= TOL[precision]
        A1, B1 = _pair(Nd, om[1:2], Jd=6, Kd=Kd, Ld=1024, precision=precision)
        assert rel_l2(A1.fft(x), B1.fft(x)) <= TOL[precision]


def test_non_finite_points_do_not_fault():
    """NaN / inf coordinates give non-finite samples, never an out-of-bounds access."""
    Nd, Kd = (16, 16), (32, 32)
    om = np.array([[np.nan, 0.0], [np.inf, 1.0], [0.3, -np.inf], [0.1, 0.2], [1e30, -1e30]])
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    x = trj.complex_gaussian(Nd, seed=15)
    y = A.fft(x)
    assert np.isfinite(y[3])
    xa = A.adj(np.ones(5))
    assert xa.shape == Nd
    torch.cuda.synchronize()


# ------------------------------------------------ kernel families / binning
BINNINGS = [
    # (method, bin_size, max_points)
    ("gm", None, 0),
    ("tiled", None, 0),
    ("tiled_atomic", None, 0),
    ("tiled_two_phase", (8, 8), 100),
    ("tiled", (8, 8), 7),        # many small subproblems
    ("tiled", (64, 16), 1000),   # anisotropic bins
    ("tiled", (200, 40), 64),    # bins larger than the grid on axis 0 (clamped to K)
]


@pytest.mark.parametrize("method,bin_size,max_points", BINNINGS)
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_binning_variants_vs_oracle(method, bin_size, max_points, precision, phasing):
    Nd, Kd = (40, 36), (80, 75)
    om = trj.golden_angle_2d(50, 80)
    A, B = _pair(Nd, om, Jd=(6, 5), Kd=Kd, Ld=1024, precision=precision, phasing=phasing, ortho=True,
                 n_shift=(7.5, -3.0))
    A.set_binning(bin_size, max_points, method)
    info = A.plan_info()
    assert info["tiled"] == method.startswith("tiled")
    if method == "tiled_two_phase":
        assert info["two_phase"]
    if method == "tiled_atomic":
        assert not info["two_phase"]
    x = trj.complex_gaussian(Nd + (2,), seed=21)
    y = A.fft(x)
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    assert rel_l2(A.adj(y), B.adj(y)) <= TOL[precision]


@pytest.mark.parametrize("max_points", [500, 250, 125, 55, 54, 18])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_dense_bin_dealt_subproblems_vs_oracle(max_points, precision):
    """One 16x16 bin holding about 1,000 points, many of them in the same few
    grid cells (deep rank levels), cut into ns = 2, 4, 8, 16, 17 and 48
    subproblems (for these seeds): even counts take the rank-rotated key
    order of the bin sort, odd ones the plain order.  Forward and adjoint against the oracle, with
    the fixed-point and the float64 scatter tile for complex64."""
    Nd, Kd = (32, 32), (64, 64)
    rng = np.random.default_rng(max_points)
    t = np.concatenate([rng.uniform(20.0, 36.0, (600, 2)),                      # spread over the bin
                        np.repeat(rng.uniform(24.0, 30.0, (8, 2)), 50, axis=0)])  # 8 cells x 50 points
    om = 2 * np.pi * t / np.asarray(Kd) - np.pi
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
    A.set_binning((16, 16), max_points, "tiled")
    x = trj.complex_gaussian(Nd + (2,), seed=max_points)
    y = A.fft(x)
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    xo = B.adj(y)
    assert rel_l2(A.adj(y), xo) <= TOL[precision]
    if precision == "single":
        A.set_scatter("float64")
        assert rel_l2(A.adj(y), xo) <= TOL[precision]


@pytest.mark.parametrize("ndim", [1, 3])
def test_tiled_1d_3d_small_grids(ndim):
    """Tiles wider than the grid (K < B + J - 1) wrap onto themselves."""
    if ndim == 1:
        Nd, Kd, om = (10,), (20,), np.random.default_rng(0).uniform(-np.pi, np.pi, (300, 1))
    else:
        Nd, Kd, om = (6, 8, 5), (12, 16, 10), trj.kooshball_3d(40, 16, seed=2)
    for precision in ("single", "double"):
        A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=512, precision=precision)
        A.set_binning(None, 5, "tiled")
        x = trj.complex_gaussian(Nd, seed=22)
        y = A.fft(x)
        assert rel_l2(y, B.fft(x)) <= TOL[precision]
        assert rel_l2(A.adj(y), B.adj(y)) <= TOL[precision]


def test_default_binning_active():
    A = NufftBase((64, 64), trj.radial_2d(16, 128), Jd=6, Kd=(128, 128), device=dev())
    info = A.plan_info()
    assert info["tiled"] and info["bin_size"][:2] == (16, 16)
    assert info["nbins"] == 64 and info["tile"] == 21 * 21


def test_tiled_rejects_tile_over_lds_budget():
    """A forced tiled method with a tile beyond the LDS budget is an error."""
    A = NufftBase((80, 72), trj.golden_angle_2d(10, 160), Jd=6, Kd=(160, 150), precision="double", device=dev())
    with pytest.raises(ValueError, match="LDS"):
        A.set_binning((200, 200), 64, "tiled")  # 165 x 155 complex128 points: 409 KiB


def test_graph_capture_plain_operator():
    """forward + adjoint captured in one HIP graph; replays equal the eager
    result, with eager work and synchronisations between replays."""
    Nd, Kd = (64, 64), (128, 128)
    om = trj.golden_angle_2d(40, 128)
    for precision in ("single", "double"):
        A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, device=dev())
        cdt = A._tcdt
        x = torch.from_numpy(trj.complex_gaussian(Nd, seed=30)).to(dev()).to(cdt).unsqueeze(-1)
        y_ref = A._forward_dev(x, 1).clone()
        xa_ref = A._adjoint_dev(y_ref, 1).clone()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            A._forward_dev(x, 1)
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            y = A._forward_dev(x, 1)
            xa = A._adjoint_dev(y, 1)
        for _ in range(3):
            g.replay()
            torch.cuda.synchronize()
            assert rel_l2(y.cpu().numpy(), y_ref.cpu().numpy()) <= 1e-6
            assert rel_l2(xa.cpu().numpy(), xa_ref.cpu().numpy()) <= 1e-6
            _ = (x * 3).abs().sum().item()


def test_kernel_timing_events():
    """Event-timed tiled launches: counted while enabled, none while disabled
    or while the stream is captured; reading resets the totals."""
    Nd, Kd = (64, 64), (128, 128)
    A = NufftBase(Nd, trj.golden_angle_2d(40, 128), Jd=6, Kd=Kd, Ld=1024, device=dev())
    x = torch.from_numpy(trj.complex_gaussian(Nd, seed=31)).to(dev()).to(A._tcdt).unsqueeze(-1)
    A._forward_dev(x, 1)
    assert A.read_kernel_timing("interp")[0] == 0
    A.set_kernel_timing(True)
    for _ in range(3):
        y = A._forward_dev(x, 1)
        A._adjoint_dev(y, 1)
    n_i, ms_i = A.read_kernel_timing("interp")
    n_s, ms_s = A.read_kernel_timing("spread")
    assert (n_i, n_s) == (3, 3)
    assert 0.0 < ms_i < 100.0 and 0.0 < ms_s < 100.0
    assert A.read_kernel_timing("spread")[0] == 0
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        A._forward_dev(x, 1)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    A.read_kernel_timing("interp")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        A._forward_dev(x, 1)
    g.replay()
    torch.cuda.synchronize()
    assert A.read_kernel_timing("interp")[0] == 0
    A.set_kernel_timing(False)
    with pytest.raises(KeyError):
        A.read_kernel_timing("fft")


def test_flush_atomic_count():
    """The scatter's flush makes one float atomic per component of each
    non-zero tile cell outside its bin's own cells, which get plain stores
    (local index J-1 <= e < B on every axis, 16x16 bins, J = 6).  A point whose
    taps start at local cell (10, 3) covers 6 x 6 cells, 6 x 4 of them its
    bin's own: 12 cells, 24 atomics per repetition.  One starting at (0, 0)
    of another bin has a single own cell: 35 cells, 70 atomics.  Counted only
    while kernel timing is on with the count asked for."""
    Nd, Kd = (64, 64), (128, 128)
    t = np.array([[28.5, 21.5], [34.5, 34.5]])  # first taps (26, 19) and (32, 32)
    A = NufftBase(Nd, t * 2 * np.pi / 128, Jd=6, Kd=Kd, Ld=1024, device=dev())
    assert A.plan_info()["bin_size"][:2] == (16, 16)
    y = torch.ones((3, 2), dtype=A._tcdt, device=dev())
    A._adjoint_dev(y, 3)
    assert A.read_flush_atomics() == 0
    A.set_kernel_timing(True)
    A._adjoint_dev(y, 3)
    assert A.read_flush_atomics() == 0
    A.read_kernel_timing("spread")
    A.set_kernel_timing(True, count_atomics=True)
    A._adjoint_dev(y, 3)
    A._adjoint_dev(y[:1].contiguous(), 1)
    assert A.read_kernel_timing("spread")[0] == 2
    assert A.read_flush_atomics() == 4 * (24 + 70)
    assert A.read_flush_atomics() == 0
    A.set_kernel_timing(False)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_grid_only_and_psf_vs_oracle(precision, phasing):
    """nufft_forward(grid_only) / nufft_adj(grid_only, return_psf)
    (_nufft.py:1302-1304, :1376-1388, :1495, :1507-1518)."""
    from mrrt.nufft_amd import nufft_adj, nufft_forward
    Nd, Kd = (24, 20), (48, 40)
    om = trj.golden_angle_2d(12, 48)
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, phasing=phasing, n_shift=(3.5, 2.0))
    g = trj.complex_gaussian((int(np.prod(Kd)), 2), seed=40)
    assert rel_l2(nufft_forward(A, g, grid_only=True), B.fft(g, grid_only=True)) <= TOL[precision]
    y = trj.complex_gaussian((om.shape[0], 2), seed=41)
    assert rel_l2(nufft_adj(A, y, grid_only=True), B.adj(y, grid_only=True)) <= TOL[precision]
    psf = nufft_adj(A, np.ones(om.shape[0]), return_psf=True)
    assert psf.shape == Kd
    assert rel_l2(psf, B.adj(np.ones(om.shape[0]), return_psf=True)) <= TOL[precision]
    # both flags: grid_only is checked first (:1507-1509 before :1517), so the
    # flat gridded samples come back, without conj(phase_after)
    both = nufft_adj(A, y, return_psf=True, grid_only=True)
    assert both.shape == (int(np.prod(Kd)), 2)
    assert rel_l2(both, B.adj(y, return_psf=True, grid_only=True)) <= TOL[precision]
    # the reference's positional order: (obj, x, copy_x, grid_only, xp) and
    # (obj, xk, copy, return_psf, grid_only, xp)
    assert rel_l2(nufft_forward(A, g, True, True, np), B.fft(g, grid_only=True)) <= TOL[precision]
    assert rel_l2(nufft_adj(A, y, True, False, True, np), B.adj(y, grid_only=True)) <= TOL[precision]


def test_two_phase_scatter_is_deterministic():
    """The two-phase scatter sums in a fixed order: repeated adjoints are
    bit-identical (the atomic flush is not, in general)."""
    Nd, Kd = (64, 64), (128, 128)
    om = trj.golden_angle_2d(60, 128)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    A.set_binning(None, 0, "tiled_two_phase")
    y = trj.complex_gaussian((om.shape[0],), seed=50).astype(np.complex64)
    ref = A.adj(y)
    for _ in range(3):
        assert np.array_equal(A.adj(y), ref)


@pytest.mark.parametrize("ndim", [1, 3])
def test_two_phase_small_grids_wrap(ndim):
    """Tiles wider than the grid cover a grid point at several offsets."""
    if ndim == 1:
        Nd, Kd, om = (10,), (20,), np.random.default_rng(1).uniform(-np.pi, np.pi, (300, 1))
    else:
        Nd, Kd, om = (6, 8, 5), (12, 16, 10), trj.kooshball_3d(40, 16, seed=4)
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=512, precision="double")
    A.set_binning(None, 7, "tiled_two_phase")
    y = trj.complex_gaussian((om.shape[0],), seed=51)
    assert rel_l2(A.adj(y), B.adj(y)) <= 1e-12
