"""Multi-coil (SENSE) encoding around the NUFFT against a per-coil loop of
the oracle:

    E x = [A (S_c x)]_c,   E^H W y = sum_c conj(S_c) A^H (W y_c)

Bars: relative L2 <= 1e-5 (complex64) / 1e-12 (complex128) against the
oracle; <E x, y> = <x, E^H y> to rounding; the Toeplitz normal operator
within the NUFFT's own accuracy (2e-3) of E^H W E.
"""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase
from mrrt.nufft_amd import trajectories as trj
from mrrt.nufft_amd.batched import NufftFrames
from mrrt.nufft_amd.mri import SenseOperator
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def _maps(C, Nd, seed):
    """Smooth complex coil sensitivities on the image grid."""
    rng = np.random.default_rng(seed)
    grids = np.meshgrid(*[np.linspace(-1, 1, n) for n in Nd], indexing="ij")
    maps = []
    for _ in range(C):
        c = rng.uniform(-1, 1, len(Nd))
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        maps.append(np.exp(-r2) * np.exp(1j * rng.uniform(0, 2 * np.pi)))
    return np.stack(maps)


def _oracle_forward(B, smaps, x):
    return np.stack([B.fft(smaps[c] * x) for c in range(smaps.shape[0])])


def _oracle_adjoint(B, smaps, y, w):
    w = np.ones(y.shape[1]) if w is None else w
    return sum(np.conj(smaps[c]) * B.adj((w * y[c]).astype(B.cdt)) for c in range(smaps.shape[0]))


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("Nd,Kd", [((48, 40), (96, 80)), ((16, 12, 10), (32, 24, 20))])
def test_sense_vs_oracle(precision, weighted, Nd, Kd):
    om = trj.golden_angle_2d(30, 96) if len(Nd) == 2 else trj.kooshball_3d(40, 32, seed=2)
    C = 5
    smaps = _maps(C, Nd, seed=1)
    w = np.random.default_rng(3).uniform(0.2, 1.0, om.shape[0]) if weighted else None
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, device=dev())
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision)
    E = SenseOperator(A, smaps, weights=w)
    x = trj.complex_gaussian(Nd, seed=4)
    y = E.forward(x)
    assert y.shape == (C, om.shape[0])
    assert rel_l2(y, _oracle_forward(B, smaps.astype(B.cdt), x.astype(B.cdt))) <= TOL[precision]
    yy = trj.complex_gaussian((C, om.shape[0]), seed=5)
    assert rel_l2(E.adjoint(yy), _oracle_adjoint(B, smaps.astype(B.cdt), yy, w)) <= TOL[precision]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_sense_adjointness(precision):
    Nd, Kd = (40, 36), (80, 72)
    om = np.random.default_rng(6).uniform(-np.pi, np.pi, (900, 2))
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, device=dev())
    E = SenseOperator(A, _maps(4, Nd, seed=7))
    x = trj.complex_gaussian(Nd, seed=8)
    y = trj.complex_gaussian((4, om.shape[0]), seed=9)
    lhs = np.vdot(E.forward(x), y)
    rhs = np.vdot(x, E.adjoint(y))
    assert abs(lhs - rhs) <= {"single": 1e-5, "double": 1e-12}[precision] * abs(lhs)


def test_sense_normal_and_toeplitz():
    Nd, Kd = (48, 48), (96, 96)
    om = trj.golden_angle_2d(40, 96)
    w = np.random.default_rng(10).uniform(0.5, 1.0, om.shape[0])
    A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev())
    E = SenseOperator(A, _maps(6, Nd, seed=11), weights=w)
    x = trj.complex_gaussian(Nd, seed=12)
    n1 = E.normal(x)
    assert rel_l2(n1, E.adjoint(E.forward(x))) <= 1e-6
    nt = E.normal(x, toeplitz=True)
    assert rel_l2(nt, n1) <= 2e-3


def test_sense_frames_vs_oracle():
    Nd, Kd = (32, 32), (64, 64)
    spokes = [7, 11, 5]
    om_all = trj.golden_angle_2d(sum(spokes), 64)
    offs = np.concatenate([[0], np.cumsum(spokes)]) * 64
    omegas = [om_all[offs[f]:offs[f + 1]] for f in range(len(spokes))]
    C = 4
    smaps = _maps(C, Nd, seed=13)
    Fr = NufftFrames(Nd, omegas, C, Jd=6, Kd=Kd, Ld=1024, device=dev())
    w = np.random.default_rng(14).uniform(0.5, 1.0, Fr.M)
    E = SenseOperator(Fr, smaps, weights=w)
    x = trj.complex_gaussian((len(spokes),) + Nd, seed=15)
    y = E.forward(x)
    yy = trj.complex_gaussian((C, Fr.M), seed=16)
    xa = E.adjoint(yy)
    for f in range(len(spokes)):
        B = O.NufftOracle(Nd, omegas[f], Jd=6, Kd=Kd, Ld=1024)
        sl = Fr.frame_slice(f)
        s64 = smaps.astype(np.complex64)
        assert rel_l2(y[:, sl], _oracle_forward(B, s64, x[f].astype(np.complex64))) <= 1e-5
        assert rel_l2(xa[f], _oracle_adjoint(B, s64, yy[:, sl], w[sl])) <= 1e-5
    with pytest.raises(ValueError):
        E.normal(x, toeplitz=True)


def test_sense_rejects_bad_shapes():
    Nd = (16, 16)
    A = NufftBase(Nd, trj.golden_angle_2d(8, 32), Jd=6, Kd=(32, 32), device=dev())
    with pytest.raises(ValueError):
        SenseOperator(A, np.ones((3, 16, 15), np.complex64))
    E = SenseOperator(A, np.ones((3,) + Nd, np.complex64))
    with pytest.raises(ValueError):
        E.forward(np.ones((16, 17)))
    with pytest.raises(ValueError):
        SenseOperator(A, np.ones((3,) + Nd), weights=np.ones(5))
