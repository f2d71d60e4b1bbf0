"""The scatter's 32-bit fixed-point LDS tile (complex64): per-subproblem
power-of-two scales bounded by the samples' magnitudes.

Checked against the oracle's adjoint (nufft_adj restated,
mrrt/nufft/_nufft.py:1459-1578) and against the float64 tile of the same
plan: at the parity bar for ordinary inputs, without overflow when every
sample sits in one grid cell (the bound is tight there), exactly scale-
covariant over the float range (the scale is a power of two), NaN samples
propagate, zeros stay zeros.
"""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase
from mrrt.nufft_amd import trajectories as trj
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def _case(ndim):
    if ndim == 2:
        Nd, Kd = (40, 48), (80, 96)
        om = trj.golden_angle_2d(60, 80)
    else:
        Nd, Kd = (24, 20, 16), (48, 40, 32)
        om = trj.kooshball_3d(400, 48, seed=4)
    return Nd, Kd, om


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_fixed_point_vs_oracle_and_float64(ndim, phasing):
    Nd, Kd, om = _case(ndim)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing=phasing, n_shift=np.asarray(Nd) / 2, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing=phasing, n_shift=np.asarray(Nd) / 2)
    rng = np.random.default_rng(ndim)
    y = (rng.standard_normal((om.shape[0], 3)) + 1j * rng.standard_normal((om.shape[0], 3))).astype(np.complex64)
    A.set_scatter("fixed32")
    assert A.plan_info()["fixed_point_scatter"]
    xf = A.adj(y)
    A.set_scatter("float64")
    assert not A.plan_info()["fixed_point_scatter"]
    xd = A.adj(y)
    xo = B.adj(y)
    assert rel_l2(xf, xo) <= 1e-5
    assert rel_l2(xd, xo) <= 1e-5
    assert rel_l2(xf, xd) <= 5e-6


def test_default_accumulation_per_dimension():
    """2-D and 3-D complex64 plans default to the fixed-point tile,
    complex128 plans to float64."""
    for ndim, precision, want in ((2, "single", True), (3, "single", True), (2, "double", False),
                                  (3, "double", False)):
        Nd, Kd, om = _case(ndim)
        A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, precision=precision, device=dev())
        assert A.plan_info()["fixed_point_scatter"] == want, (ndim, precision)


@pytest.mark.parametrize("ndim", [2, 3])
def test_fixed_point_all_samples_in_one_cell(ndim):
    """5,000 samples of one grid cell with equal values: every cell sum is at
    its bound (the tile must not overflow), and identical taps round
    identically, so their rounding adds coherently (the worst case of the
    fixed-point error: about 5e-6 by the bound, within the parity bar)."""
    Nd, Kd, _ = _case(ndim)
    om = np.tile(np.full((1, ndim), 0.1234), (5000, 1)) + 1e-9 * np.arange(5000)[:, None]
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, device=dev())
    y = np.full(5000, 1.0 + 0.5j, np.complex64)
    A.set_scatter("fixed32")
    xf = A.adj(y)
    A.set_scatter("float64")
    xd = A.adj(y)
    assert np.all(np.isfinite(xf))
    assert rel_l2(xf, xd) <= 1e-5


def test_fixed_point_range_nan_and_zero():
    """Inputs scaled by 2^-120 .. 2^100 keep the fixed-point tile's accuracy
    (its scale is a power of two chosen per subproblem): the adjoint equals
    the float64 tile's at every scale; NaN samples propagate; zeros stay
    zeros."""
    Nd, Kd, om = _case(3)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, device=dev())
    rng = np.random.default_rng(9)
    y = (rng.standard_normal(om.shape[0]) + 1j * rng.standard_normal(om.shape[0])).astype(np.complex64)
    for s in (1.0, 2.0 ** -100, 2.0 ** 100, 2.0 ** -120):
        ys = (y * s).astype(np.complex64)
        A.set_scatter("fixed32")
        xf = A.adj(ys)
        A.set_scatter("float64")
        xd = A.adj(ys)
        # in complex128: float32 squares of these values would underflow
        assert rel_l2(xf.astype(np.complex128), xd.astype(np.complex128)) <= 5e-6, s
    A.set_scatter("fixed32")
    assert np.count_nonzero(A.adj(np.zeros_like(y))) == 0
    yn = y.copy()
    yn[17] = np.nan
    assert np.isnan(A.adj(yn)).any()


@pytest.mark.parametrize("method", ["auto", "tiled_two_phase"])
def test_two_repetitions_per_pass_vs_oracle(method):
    """Subproblems larger than the workgroup (512 points, 256 threads) with
    several repetitions: the scatter accumulates two repetitions per pass
    (two fixed-point tiles of 32x32 bins fit the chunk budget), and an odd
    count leaves a last pass of one.  Each repetition has its own scale:
    the second is 2^40 times the first."""
    Nd, Kd = (64, 64), (128, 128)
    om = trj.golden_angle_2d(60, 128)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024)
    A.set_binning((32, 32), 512, method)
    assert A.plan_info()["fixed_point_scatter"]
    rng = np.random.default_rng(21)
    y = (rng.standard_normal((om.shape[0], 3)) + 1j * rng.standard_normal((om.shape[0], 3))).astype(np.complex64)
    y[:, 1] *= np.float32(2.0 ** 40)
    x = A.adj(y)
    xo = B.adj(y)
    for r in range(3):
        assert rel_l2(x[..., r], xo[..., r]) <= 1e-5, r
    A.set_scatter("float64")
    xd = A.adj(y)
    for r in range(3):
        assert rel_l2(x[..., r].astype(np.complex128), xd[..., r].astype(np.complex128)) <= 5e-6, r


@pytest.mark.parametrize("max_points,threads", [(256, 256), (512, 256), (200, 128)])
@pytest.mark.parametrize("nreps", [1, 3])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_points_kept_across_the_scale_sum_vs_oracle(max_points, threads, nreps, phasing):
    """Subproblems of at most one point per thread (P = threads = 256): the
    scatter computes each point's coefficients once, before the scale's
    block sum, and keeps the sample that sum read for the scatter, for one
    repetition and for several (each with its own scale: the second is
    2^40 times the first).  At most two points per thread (P 512 / 256
    threads, P 200 / 128 threads) and one repetition: the indices and
    samples the sum read are kept for the scatter (several repetitions take
    the two-per-pass kernel).  A NaN sample still makes its tile NaN."""
    Nd, Kd, om = _case(2)
    shift = np.asarray(Nd) / 2
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing=phasing, n_shift=shift, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing=phasing, n_shift=shift)
    A.set_binning((16, 16), max_points, "tiled", threads)
    info = A.plan_info()
    assert info["fixed_point_scatter"] and info["max_points"] == max_points and info["threads"] == threads
    rng = np.random.default_rng(30 + nreps)
    shape = (om.shape[0],) if nreps == 1 else (om.shape[0], nreps)
    y = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    if nreps > 1:
        y[:, 1] *= np.float32(2.0 ** 40)
    x = A.adj(y)
    xo = B.adj(y)
    xs, xos = (x[..., None], xo[..., None]) if nreps == 1 else (x, xo)
    for r in range(nreps):
        assert rel_l2(xs[..., r], xos[..., r]) <= 1e-5, r
    A.set_scatter("float64")
    xd = A.adj(y)
    xds = xd[..., None] if nreps == 1 else xd
    for r in range(nreps):
        assert rel_l2(xds[..., r], xos[..., r]) <= 1e-5, r
    A.set_scatter("fixed32")
    yn = y.copy()
    yn[17] = np.nan
    assert np.isnan(A.adj(yn)).any()


@pytest.mark.parametrize("bins,max_points,threads,method", [
    ((16, 16, 16), 512, 256, "tiled"),             # one repetition per pass
    ((16, 16, 16), 512, 256, "tiled_two_phase"),   # tile slots: every cell unpacked
    ((8, 8, 8), 512, 256, "tiled"),                # two 13^3 tiles per pass
    ((8, 8, 8), 128, 128, "tiled"),                # one point per thread: coefficient cache
])
def test_packed_3d_tile_vs_oracle(bins, max_points, threads, method):
    """3-D complex64 tiles pack each cell's two fixed-point components into
    one 64-bit word (im * 2^32 + re).  Negative and positive parts of both
    components, one repetition 2^40 times another, through each scatter
    variant, against the oracle and the float64 tile."""
    Nd, Kd, om = _case(3)
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing="complex", n_shift=np.asarray(Nd) / 2, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=512, phasing="complex", n_shift=np.asarray(Nd) / 2)
    A.set_binning(bins, max_points, method, threads)
    assert A.plan_info()["fixed_point_scatter"]
    rng = np.random.default_rng(77)
    y = (rng.standard_normal((om.shape[0], 2)) + 1j * rng.standard_normal((om.shape[0], 2))).astype(np.complex64)
    y[:, 1] *= np.float32(2.0 ** 40)
    x = A.adj(y)
    xo = B.adj(y)
    for r in range(2):
        assert rel_l2(x[..., r], xo[..., r]) <= 1e-5, r
    A.set_scatter("float64")
    xd = A.adj(y)
    for r in range(2):
        assert rel_l2(x[..., r].astype(np.complex128), xd[..., r].astype(np.complex128)) <= 5e-6, r
