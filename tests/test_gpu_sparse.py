"""Sparse-matrix mode (NufftBase(mode='sparse'), mrrt/nufft/_nufft.py:751-877,
used at :1383-1384 and :1505) on the GPU against the oracle's restatement
and against the exact DTFT at the reference's own tolerances
(mrrt/nufft/tests/test_nufft.py:60-324, which run every case in both modes).

Bars: relative L2 <= 1e-5 (complex64) / 1e-12 (complex128) against the
oracle; the matrix itself (read out through the bare product) equals the
oracle's float32 matrix exactly for the real kernel.
"""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase, nufft_adj, nufft_forward
from mrrt.nufft_amd import trajectories as trj
from oracle import nufft_oracle as O
from tests.cases import perturbed_gridpoints, rel_l2

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def _pair(Nd, om, **kw):
    return NufftBase(Nd, om, mode="sparse", device=dev(), **kw), O.NufftOracle(Nd, om, mode="sparse", **kw)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("Nd,Kd,Jd,n_shift,ortho", [
    ((64,), (128,), 6, (5.5,), False),
    ((32, 24), (64, 49), (6, 7), (3.3, -2.0), False),
    ((48, 40), (96, 80), 6, None, True),
    ((16, 12, 10), (32, 24, 20), (6, 5, 4), (1.0, 2.0, -3.0), False),
    # power-of-two grids: phase_before in the fused FFT passes
    ((48, 40), (64, 128), 6, (2.0, -1.0), True),
    ((12, 10, 8), (32, 16, 16), 6, None, False),
])
def test_sparse_operator_vs_oracle(precision, phasing, Nd, Kd, Jd, n_shift, ortho):
    if len(Nd) == 2:
        om = trj.golden_angle_2d(40, 2 * Nd[0])           # dense centre: columns longer than one piece
    else:
        om = np.random.default_rng(2).uniform(-np.pi, np.pi, (700, len(Nd)))
    A, B = _pair(Nd, om, Jd=Jd, Kd=Kd, Ld=1024, n_shift=n_shift, ortho=ortho, precision=precision,
                 phasing=phasing)
    info = A.plan_info()
    assert info["sparse_nnz"] == om.shape[0] * int(np.prod(np.broadcast_to(Jd, (len(Nd),))))
    if len(Nd) == 2:
        assert info["sparse_pieces"] > int(np.prod(Kd))   # some columns are split
    assert A.sparse_format == "CSC"
    x = trj.complex_gaussian(Nd + (3,), seed=1)
    y = A.fft(x)
    assert y.dtype == B.cdt
    assert rel_l2(y, B.fft(x)) <= TOL[precision]
    yb = trj.complex_gaussian((om.shape[0], 6), seed=2)   # > 4 repetitions: two passes of the forward rows
    assert rel_l2(A.adj(yb), B.adj(yb)) <= TOL[precision]
    assert rel_l2(A.fft(x[..., 0]), B.fft(x[..., 0])) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_sparse_matrix_entries(precision, phasing):
    """P read column by column through the bare product (grid_only forward
    of identity grids) equals the oracle's P."""
    Nd, Kd = (6, 5), (12, 10)
    om = np.random.default_rng(3).uniform(-np.pi, np.pi, (40, 2))
    A, B = _pair(Nd, om, Jd=(4, 3), Kd=Kd, Ld=512, n_shift=(1.5, 2.5), precision=precision, phasing=phasing)
    nk = int(np.prod(Kd))
    P = nufft_forward(A, np.eye(nk, dtype=B.cdt), grid_only=True)
    ref = B.p.toarray()
    if phasing == "real":
        np.testing.assert_array_equal(P, ref)
    else:
        assert rel_l2(P, ref) <= TOL[precision]
    # and P^H through the grid_only adjoint, which applies conj(phase_after)
    # first (_nufft.py:1495 precedes the grid_only return at :1507-1509)
    eye = np.eye(om.shape[0], dtype=B.cdt)
    PH = nufft_adj(A, eye, grid_only=True)
    assert rel_l2(PH, B.adj(eye, grid_only=True)) <= TOL[precision]
    pa = np.ones(om.shape[0]) if B.phase_after is None else B.phase_after
    assert rel_l2(PH, ref.conj().T * np.conj(pa)[None, :]) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_sparse_grid_only_and_psf_vs_oracle(precision, phasing):
    Nd, Kd = (24, 20), (48, 40)
    om = trj.golden_angle_2d(12, 48)
    A, B = _pair(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, phasing=phasing, n_shift=(3.5, 2.0))
    g = trj.complex_gaussian((int(np.prod(Kd)), 2), seed=40)
    assert rel_l2(nufft_forward(A, g, grid_only=True), B.fft(g, grid_only=True)) <= TOL[precision]
    y = trj.complex_gaussian((om.shape[0], 2), seed=41)
    assert rel_l2(nufft_adj(A, y, grid_only=True), B.adj(y, grid_only=True)) <= TOL[precision]
    psf = nufft_adj(A, np.ones(om.shape[0]), return_psf=True)
    assert rel_l2(psf, B.adj(np.ones(om.shape[0]), return_psf=True)) <= TOL[precision]


def test_sparse_phase_after_from_cast_omega():
    """Single precision: phase_after from the float32 omega (_nufft.py:768-769),
    and the operator's output follows it."""
    Nd = (16, 16)
    om = perturbed_gridpoints(Nd)
    A = NufftBase(Nd, om, Jd=6, Kd=(32, 32), mode="sparse", device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=(32, 32), mode="sparse")
    np.testing.assert_array_equal(A.phase_after.cpu().numpy(), B.phase_after)
    x = trj.complex_gaussian(Nd, seed=3)
    assert rel_l2(A.fft(x), B.fft(x)) <= 1e-5


def test_sparse_adjoint_bit_reproducible_and_close_to_table():
    Nd, Kd = (64, 64), (128, 128)
    om = trj.golden_angle_2d(60, 128)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, mode="sparse", device=dev())
    T = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, mode="table", device=dev())
    y = torch.from_numpy(trj.complex_gaussian((2, A.M), seed=4)).to(dev()).to(A._tcdt)
    a1 = A._adjoint_dev(y, 2).clone()
    a2 = A._adjoint_dev(y, 2).clone()
    assert torch.equal(a1, a2)
    # exact-kernel matrix vs the L = 1024 table: the same NUFFT up to the
    # table's interpolation error
    assert rel_l2(a1.cpu().numpy(), T._adjoint_dev(y, 2).cpu().numpy()) <= 1e-3
    x = torch.from_numpy(trj.complex_gaussian(Nd + (1,), seed=5)).to(dev()).to(A._tcdt)
    assert rel_l2(A._forward_dev(x, 1).cpu().numpy(), T._forward_dev(x, 1).cpu().numpy()) <= 1e-3


def test_sparse_set_points_rebuilds():
    Nd, Kd = (20, 24), (40, 48)
    om1 = trj.golden_angle_2d(10, 40)
    om2 = np.random.default_rng(6).uniform(-np.pi, np.pi, (333, 2))
    A = NufftBase(Nd, om1, Jd=6, Kd=Kd, mode="sparse", device=dev())
    A.set_points(om2)
    B = O.NufftOracle(Nd, om2, Jd=6, Kd=Kd, mode="sparse")
    assert A.plan_info()["sparse_nnz"] == 333 * 36
    x = trj.complex_gaussian(Nd, seed=7)
    assert rel_l2(A.fft(x), B.fft(x)) <= 1e-5
    y = trj.complex_gaussian((333,), seed=8)
    assert rel_l2(A.adj(y), B.adj(y)) <= 1e-5


def test_sparse_clear_returns_to_table():
    from mrrt.nufft_amd import _lib
    Nd, Kd = (20, 24), (40, 48)
    # float32 points: the sparse and table phase_after then agree (both from
    # the float32 omega)
    om = trj.golden_angle_2d(10, 40).astype(np.float32)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, mode="sparse", device=dev())
    _lib.check(A._lib.mi355nufft_clear_sparse(A._plan))
    assert A.plan_info()["sparse_nnz"] is None
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, mode="table")
    x = trj.complex_gaussian(Nd, seed=9)
    assert rel_l2(A.fft(x), B.fft(x)) <= 1e-5
    with pytest.raises(RuntimeError):
        _lib.check(A._lib.mi355nufft_sparse_apply(A._plan, 0, None, None, 1))


def test_sparse_rejects_trajectory_gradient():
    from mrrt.nufft_amd.autograd import nufft
    Nd = (16, 16)
    om = trj.golden_angle_2d(8, 32)
    A = NufftBase(Nd, om, Jd=6, Kd=(32, 32), mode="sparse", device=dev())
    x = torch.from_numpy(trj.complex_gaussian((1,) + Nd, seed=10)).to(dev()).to(A._tcdt)
    omega = torch.from_numpy(om).to(dev()).requires_grad_(True)
    with pytest.raises(RuntimeError):
        nufft(A, x, omega).abs().sum().backward()


# ---------------------------------- reference test cases in sparse mode (DTFT)
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_sparse_nufft_adj_kat(phasing):
    """mrrt/nufft/tests/test_nufft.py:60-96 (sparse mode)."""
    o1 = 2 * np.pi * np.array([0.0, 0.1, 0.3, 0.4, 0.7, 0.9])
    om = np.stack((o1, o1[::-1].copy()), axis=-1)
    st = NufftBase(omega=om, Nd=(4, 8), Jd=[8, 8], Kd=(8, 16), Ld=1024, n_shift=[2.7, 3.1], phasing=phasing,
                   mode="sparse", device=dev())
    data = np.arange(1, 7) ** 2
    xd = O.dtft_adj(data, om, (4, 8), [2.7, 3.1]).reshape((4, 8), order="F")
    xn = st.adj(np.stack((data,) * 3, axis=-1))
    np.testing.assert_array_almost_equal(xd, xn[:, :, -1], decimal=4)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("order", ["F", "C"])
def test_sparse_nufft_1d_vs_dtft(precision, phasing, order):
    """mrrt/nufft/tests/test_nufft.py:99-169 (sparse mode)."""
    Nd, n_shift = 64, 32
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=6, Kd=128, n_shift=n_shift, mode="sparse", Ld=1024, precision=precision,
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
def test_sparse_nufft_2d_vs_dtft(precision, phasing, Kd, Jd, order):
    """mrrt/nufft/tests/test_nufft.py:172-244 (sparse mode)."""
    Nd = (16, 16)
    n_shift = np.asarray(Nd) / 2
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=Jd, Kd=Kd, n_shift=n_shift, mode="sparse", Ld=1024, precision=precision,
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
def test_sparse_nufft_3d_vs_dtft(precision, phasing, order):
    """mrrt/nufft/tests/test_nufft.py:247-324 (sparse mode)."""
    Nd, Kd = [8] * 3, [16] * 3
    n_shift = np.asarray(Nd) / 2
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = NufftBase(omega=om, Nd=Nd, Jd=[6] * 3, Kd=Kd, n_shift=n_shift, mode="sparse", Ld=1024,
                  precision=precision, phasing=phasing, order=order, device=dev())
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
def test_empty_points(mode):
    """No sample points: the forward returns no samples; the device adjoint
    returns a zero image.  The array-level adjoint raises like the
    reference's size check (xk.size % obj.M, _nufft.py:1483)."""
    Nd = (16, 12)
    A = NufftBase(Nd, np.zeros((0, 2)), Jd=6, Kd=(32, 24), mode=mode, device=dev())
    assert A.M == 0
    y = A.fft(trj.complex_gaussian(Nd, seed=50))
    assert y.shape == (0,)
    xa = A._adjoint_dev(torch.zeros((1, 0), dtype=A._tcdt, device=dev()), 1)
    assert tuple(xa.shape) == Nd + (1,) and not torch.any(xa != 0)
    with pytest.raises(ZeroDivisionError):
        A.adj(np.zeros((0,), np.complex64))
