"""Differentiable NUFFT: gradients with respect to the image and the sampling
points, checked against FP64 central finite differences (step 1e-6) and the
reference adjoint (the oracle restatement of nufft_adj)."""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase
from mrrt.nufft_amd import trajectories as trj
from mrrt.nufft_amd.autograd import nufft, nufft_adjoint
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

H = 1e-6


def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


def _crossing_mask(op, omega, h):
    """True where omega +- h moves some tap of that axis into another table
    cell (the table interpolant has a kink there, so FD sees a secant)."""
    om = np.asarray(omega, dtype=np.float64)
    mask = np.zeros(om.shape, bool)
    for d in range(om.shape[1]):
        K, J, L = op.Kd[d], op.Jd[d], op.Ld
        cells = []
        for sgn in (-1, 1):
            t = (om[:, d] + sgn * h) / (2 * np.pi / K)
            k0 = 1 + np.floor(t - J / 2.0)
            p = (t[:, None] - (k0[:, None] + np.arange(J))) * L
            cells.append(np.floor(p) + 1e6 * k0[:, None])
        mask[:, d] = np.any(cells[0] != cells[1], axis=1)
    return mask


def _loss_fn(op, y_target):
    def loss(x, omega):
        y = nufft(op, x, omega)
        return ((y - y_target).abs() ** 2).sum()
    return loss


@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("ndim", [2, 3])
def test_grads_fp64_vs_finite_differences(phasing, ndim):
    torch.manual_seed(0)
    if ndim == 2:
        Nd, Kd = (16, 12), (32, 24)
        om = trj.golden_angle_2d(8, 32) * 0.9
    else:
        Nd, Kd = (8, 6, 8), (16, 12, 16)
        om = trj.kooshball_3d(12, 16, seed=3) * 0.9
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision="double", phasing=phasing, n_shift=np.asarray(Nd) / 2 + 0.3,
                   device=dev())
    x = torch.from_numpy(trj.complex_gaussian((2,) + Nd, seed=1)).to(dev()).requires_grad_(True)
    omega = torch.from_numpy(om).to(dev()).requires_grad_(True)
    y_t = torch.from_numpy(trj.complex_gaussian((2, om.shape[0]), seed=2)).to(dev())
    loss = _loss_fn(op, y_t)
    L0 = loss(x, omega)
    gx, gw = torch.autograd.grad(L0, (x, omega))
    # image gradient == 2 A^H (A x - y): the reference adjoint (oracle)
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision="double", phasing=phasing,
                      n_shift=np.asarray(Nd) / 2 + 0.3)
    xn = x.detach().cpu().numpy()
    r = np.stack([B.fft(xn[c]) for c in range(2)]) - y_t.cpu().numpy()
    ref_gx = np.stack([B.adj(2 * r[c]) for c in range(2)])
    assert rel_l2(gx.cpu().numpy(), ref_gx) <= 1e-12
    # image gradient vs FD along a random complex direction
    v = torch.from_numpy(trj.complex_gaussian((2,) + Nd, seed=3)).to(dev())
    with torch.no_grad():
        fd = (loss(x + H * v, omega) - loss(x - H * v, omega)) / (2 * H)
    an = torch.real(torch.sum(torch.conj(gx) * v))
    assert abs(fd.item() - an.item()) <= 1e-6 * abs(an.item())
    # point gradient vs FD, entry by entry, where no tap crosses a table cell
    gwn = gw.cpu().numpy()
    cross = _crossing_mask(op, om, H)
    rng = np.random.default_rng(4)
    checked = 0
    for m in rng.choice(om.shape[0], 12, replace=False):
        for d in range(ndim):
            if cross[m, d]:
                continue
            e = torch.zeros_like(omega)
            e[m, d] = H
            with torch.no_grad():
                fd = (loss(x, omega + e) - loss(x, omega - e)).item() / (2 * H)
            assert abs(fd - gwn[m, d]) <= 1e-5 * max(abs(gwn[m, d]), 1e-3 * np.abs(gwn).max()), (m, d, fd, gwn[m, d])
            checked += 1
    assert checked >= 6 * ndim
    assert cross.mean() < 0.2


def test_c64_gradients_vs_fp64_fd_spiral():
    """Multi-coil spiral with y from a perturbed trajectory.  The complex128
    operator's gradients are checked against FP64 central finite
    differences; the complex64 operator's gradients against those
    (relative L2), which carries the FD check to complex64."""
    Nd, Kd = (48, 48), (96, 96)
    om = trj.spiral_2d(4, 256, seed=0)
    kw = dict(Jd=6, Kd=Kd, Ld=1024, device=dev())
    op32 = NufftBase(Nd, om, precision="single", **kw)
    op64 = NufftBase(Nd, om, precision="double", **kw)
    x64 = torch.from_numpy(trj.complex_gaussian((4,) + Nd, seed=5)).to(dev())
    om64 = torch.from_numpy(om).to(dev())
    om_pert = torch.from_numpy(om + 1e-3 * np.random.default_rng(6).standard_normal(om.shape)).to(dev())
    with torch.no_grad():
        y_t = nufft(op64, x64, om_pert)
    # FP64 gradients vs FP64 finite differences
    xg = x64.clone().requires_grad_(True)
    wg = om64.clone().requires_grad_(True)
    loss64 = _loss_fn(op64, y_t)
    gx64, gw64 = torch.autograd.grad(loss64(xg, wg), (xg, wg))
    v = torch.from_numpy(trj.complex_gaussian((4,) + Nd, seed=7)).to(dev())
    with torch.no_grad():
        fd = ((loss64(x64 + H * v, om64) - loss64(x64 - H * v, om64)) / (2 * H)).item()
    an = torch.real(torch.sum(torch.conj(gx64) * v)).item()
    assert abs(fd - an) <= 1e-6 * abs(fd)
    cross = _crossing_mask(op64, om, H)
    gwn = gw64.cpu().numpy()
    scale = np.abs(gwn).max()
    checked = 0
    for m in np.random.default_rng(8).choice(om.shape[0], 16, replace=False):
        for d in range(2):
            if cross[m, d]:
                continue
            e = torch.zeros_like(om64)
            e[m, d] = H
            with torch.no_grad():
                fdv = (loss64(x64, om64 + e) - loss64(x64, om64 - e)).item() / (2 * H)
            assert abs(fdv - gwn[m, d]) <= 1e-5 * max(abs(gwn[m, d]), 1e-3 * scale), (m, d, fdv, gwn[m, d])
            checked += 1
    assert checked >= 16
    # complex64 operator vs the FD-checked FP64 gradients.  The complex64
    # operator differs from the complex128 one by ~1e-5 relative (it keeps the
    # reference's float32 phase_before, _nufft.py:703-715); the loss residual
    # r = A x - y is a small fraction of y, so that difference is amplified by
    # kappa = |y| / |r| in the gradients.
    x32 = x64.to(torch.complex64).requires_grad_(True)
    w32 = om64.to(torch.float32).requires_grad_(True)
    L32 = ((nufft(op32, x32, w32) - y_t.to(torch.complex64)).abs() ** 2).sum()
    gx32, gw32 = torch.autograd.grad(L32, (x32, w32))
    with torch.no_grad():
        y64 = nufft(op64, x64, om64)
        kappa = (torch.linalg.vector_norm(y64) / torch.linalg.vector_norm(y64 - y_t)).item()
    assert 1 < kappa < 1e3
    assert rel_l2(gx32.cpu().numpy(), gx64.cpu().numpy()) <= 5e-5 * kappa
    assert rel_l2(gw32.double().cpu().numpy(), gwn) <= 2e-4 * kappa


def test_adjoint_autograd():
    """x = A^H y: d/dy is the forward transform, d/domega vs FD."""
    Nd, Kd = (16, 16), (32, 32)
    om = trj.radial_2d(6, 32) * 0.95
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision="double", device=dev())
    y = torch.from_numpy(trj.complex_gaussian((om.shape[0],), seed=9)).to(dev()).requires_grad_(True)
    omega = torch.from_numpy(om).to(dev()).requires_grad_(True)
    target = torch.from_numpy(trj.complex_gaussian(Nd, seed=10)).to(dev())

    def loss(y_, om_):
        return ((nufft_adjoint(op, y_, om_) - target).abs() ** 2).sum()

    L0 = loss(y, omega)
    gy, gw = torch.autograd.grad(L0, (y, omega))
    with torch.no_grad():
        xa = nufft_adjoint(op, y, omega)
        ref = nufft(op, 2 * (xa - target), omega)
    assert rel_l2(gy.cpu().numpy(), ref.cpu().numpy()) <= 1e-12
    cross = _crossing_mask(op, om, H)
    gwn = gw.cpu().numpy()
    for m in range(0, om.shape[0], 17):
        for d in range(2):
            if cross[m, d]:
                continue
            e = torch.zeros_like(omega)
            e[m, d] = H
            with torch.no_grad():
                fd = (loss(y, omega + e) - loss(y, omega - e)).item() / (2 * H)
            assert abs(fd - gwn[m, d]) <= 1e-5 * max(abs(gwn[m, d]), 1e-3 * np.abs(gwn).max())


def test_points_update_in_place_is_seen():
    """An in-place update of omega (optimizer step) re-sorts the points."""
    Nd, Kd = (16, 16), (32, 32)
    om = trj.golden_angle_2d(8, 32)
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision="double", device=dev())
    x = torch.from_numpy(trj.complex_gaussian(Nd, seed=11)).to(dev())
    omega = torch.from_numpy(om).to(dev())
    y1 = nufft(op, x, omega)
    omega.mul_(0.5)
    y2 = nufft(op, x, omega)
    B = O.NufftOracle(Nd, om * 0.5, Jd=6, Kd=Kd, Ld=1024, precision="double")
    assert rel_l2(y2.cpu().numpy(), B.fft(x.cpu().numpy())) <= 1e-12
    assert rel_l2(y1.cpu().numpy(), y2.cpu().numpy()) > 1e-3


@pytest.mark.parametrize("ndim", [2, 3])
def test_point_gradient_is_sum_over_coils(ndim):
    """The trajectory gradient of an 11-coil batch (repetitions staged 8 per
    pass, then 3) equals the sum of the 11 single-coil gradients."""
    if ndim == 2:
        Nd, Kd = (16, 12), (32, 24)
        om = trj.golden_angle_2d(10, 32) * 0.9
    else:
        Nd, Kd = (8, 6, 8), (16, 12, 16)
        om = trj.kooshball_3d(12, 16, seed=5) * 0.9
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision="double", device=dev())
    C = 11
    x = torch.from_numpy(trj.complex_gaussian((C,) + Nd, seed=6)).to(dev())
    y_t = torch.from_numpy(trj.complex_gaussian((C, om.shape[0]), seed=7)).to(dev())
    omega = torch.from_numpy(om).to(dev()).requires_grad_(True)
    (gw_all,) = torch.autograd.grad(_loss_fn(op, y_t)(x, omega), (omega,))
    gw_sum = torch.zeros_like(gw_all)
    for c in range(C):
        (g,) = torch.autograd.grad(_loss_fn(op, y_t[c:c + 1])(x[c:c + 1], omega), (omega,))
        gw_sum += g
    assert rel_l2(gw_all.cpu().numpy(), gw_sum.cpu().numpy()) <= 1e-12
