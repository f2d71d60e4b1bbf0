"""Batched coils x frames with uneven per-frame trajectories, one plan, one
launch per stage; checked per frame against the oracle, and graph replays
against the eager transform."""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import trajectories as trj
from mrrt.nufft_amd.batched import NufftFrames
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_frames_vs_oracle_per_frame(precision, phasing):
    Nd, Kd = (32, 32), (64, 64)
    frames = trj.golden_angle_frames(4, 64, smin=3, smax=9, seed=1)
    frames.insert(2, np.zeros((0, 2)))            # an empty frame
    C = 3
    B = NufftFrames(Nd, frames, C, Jd=6, Kd=Kd, Ld=1024, precision=precision, phasing=phasing,
                    n_shift=(16.5, 15.0), device=dev())
    assert B.op.plan_info()["nframes"] == len(frames)
    cdt = torch.complex64 if precision == "single" else torch.complex128
    x = torch.from_numpy(trj.complex_gaussian((len(frames), C) + Nd, seed=2)).to(dev()).to(cdt)
    y = B.forward(x)
    assert tuple(y.shape) == (C, B.M)
    yt = torch.from_numpy(trj.complex_gaussian((C, B.M), seed=3)).to(dev()).to(cdt)
    xa = B.adjoint(yt.contiguous())
    xn, ynp, ytn, xan = x.cpu().numpy(), y.cpu().numpy(), yt.cpu().numpy(), xa.cpu().numpy()
    for f, om in enumerate(frames):
        sl = B.frame_slice(f)
        if om.shape[0] == 0:
            assert np.all(xan[f] == 0)
            continue
        R = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, Ld=1024, precision=precision, phasing=phasing, n_shift=(16.5, 15.0))
        ref_y = np.stack([R.fft(xn[f, c]) for c in range(C)])
        assert rel_l2(ynp[:, sl], ref_y) <= TOL[precision]
        ref_x = np.stack([R.adj(ytn[c, sl]) for c in range(C)])
        assert rel_l2(xan[f], ref_x) <= TOL[precision]


@pytest.mark.parametrize("Nd,Kd", [((48, 48), (96, 96)), ((64, 64), (128, 128))])
def test_frames_graph_replays_match_eager(Nd, Kd):
    """Replay the captured forward/adjoint several times with eager work and
    synchronisations in between; every replay equals the eager result.  Kd
    96^2 runs rocFFT with the re-zeroing crop, 128^2 the fused FFT passes,
    whose inverse passes leave the plan's adjoint grid zero (the captured
    adjoint skips clearing it): an eager adjoint of zero samples after the
    replays must return exact zeros."""
    frames = trj.golden_angle_frames(5, Kd[0], smin=4, smax=12, seed=4)
    C = 2
    B = NufftFrames(Nd, frames, C, Jd=6, Kd=Kd, Ld=1024, device=dev())
    x = torch.from_numpy(trj.complex_gaussian((5, C) + Nd, seed=5)).to(dev()).to(torch.complex64)
    y = torch.empty((C, B.M), dtype=torch.complex64, device=dev())
    xa = torch.empty((5, C) + Nd, dtype=torch.complex64, device=dev())
    y_ref = B.forward(x).clone()
    xa_ref = B.adjoint(y_ref).clone()
    B.capture(x, y, xa)
    for it in range(4):
        y.zero_()
        xa.zero_()
        B.replay_forward()
        torch.cuda.synchronize()
        assert torch.equal(y, y_ref) or rel_l2(y.cpu().numpy(), y_ref.cpu().numpy()) <= 1e-6
        B.replay_adjoint()
        torch.cuda.synchronize()
        assert rel_l2(xa.cpu().numpy(), xa_ref.cpu().numpy()) <= 1e-6
        # an eager adjoint between replays shares the plan's adjoint grid,
        # which both leave all zero
        xe = B.adjoint(y_ref)
        assert rel_l2(xe.cpu().numpy(), xa_ref.cpu().numpy()) <= 1e-6
        _ = (x * 2).sum().item()      # eager work between replays
    B.replay_forward()
    B.replay_adjoint()
    torch.cuda.synchronize()
    z = B.adjoint(torch.zeros_like(y_ref))
    assert torch.count_nonzero(z).item() == 0


def test_frames_match_separate_plans():
    """One frames plan == one plain operator per frame (same kernels, batched)."""
    from mrrt.nufft_amd import NufftBase
    Nd, Kd = (32, 40), (64, 80)
    frames = trj.golden_angle_frames(3, 64, smin=5, smax=7, seed=6)
    C = 2
    B = NufftFrames(Nd, frames, C, Jd=6, Kd=Kd, Ld=512, device=dev(), precision="double")
    x = torch.from_numpy(trj.complex_gaussian((3, C) + Nd, seed=7)).to(dev())
    y = B.forward(x).cpu().numpy()
    for f, om in enumerate(frames):
        A = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=512, device=dev(), precision="double", order="C")
        yf = A.fft(x[f].cpu().numpy())
        assert rel_l2(y[:, B.frame_slice(f)], yf) <= 1e-13
