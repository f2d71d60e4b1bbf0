"""Autotuner: the chosen binning / FFT leave the transforms unchanged (to the
precision's bar) and the cache round-trips."""

import json

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase, autotune
from mrrt.nufft_amd import trajectories as trj
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


def dev():
    assert torch.cuda.is_available()
    return torch.device("cuda", 0)


@pytest.mark.parametrize("precision,tol", [("single", 1e-5), ("double", 1e-12)])
def test_tune_keeps_results_and_caches(tmp_path, monkeypatch, precision, tol):
    monkeypatch.setenv("MI355NUFFT_TUNE_CACHE", str(tmp_path / "tune.json"))
    Nd, Kd = (64, 64), (128, 128)
    A = NufftBase(Nd, trj.golden_angle_2d(60, 128), Jd=6, Kd=Kd, Ld=1024, precision=precision, device=dev())
    x = torch.from_numpy(trj.complex_gaussian(Nd + (2,), seed=1)).to(dev()).to(A._tcdt)
    y0 = A._forward_dev(x, 2).clone()
    x0 = A._adjoint_dev(y0, 2).clone()
    cands = [((16, 16), 256, 256), ((32, 32), 128, 128)]
    b, p, timings = autotune.tune(A, nreps=2, cands=cands)
    assert (tuple(b), p) in [(c[0], c[1]) for c in cands] and len(timings) == 2
    y1 = A._forward_dev(x, 2)
    x1 = A._adjoint_dev(y1, 2)
    assert rel_l2(y1.cpu().numpy(), y0.cpu().numpy()) <= tol
    assert rel_l2(x1.cpu().numpy(), x0.cpu().numpy()) <= tol
    table = json.load(open(tmp_path / "tune.json"))
    (entry,) = table.values()
    assert entry["bin_size"] == list(b) and entry["fft"] in ("auto", "rocfft")
    B = NufftBase(Nd, trj.golden_angle_2d(60, 128), Jd=6, Kd=Kd, Ld=1024, precision=precision, device=dev())
    assert autotune.tune(B, nreps=2)[:2] == (tuple(b), p)  # from the cache
    assert B.plan_info()["bin_size"][:2] == tuple(b)


def test_tune_fft_picks_a_path_and_sets_it():
    Nd, Kd = (32, 32), (64, 64)
    A = NufftBase(Nd, trj.golden_angle_2d(30, 64), Jd=6, Kd=Kd, Ld=1024, device=dev())
    choice, ms = autotune.tune_fft(A, nreps=1, pairs=3, rounds=2)
    assert choice in ("auto", "rocfft") and set(ms) == {"auto", "rocfft"}
    assert A.plan_info()["fused_fft"] == (choice == "auto")
    C = NufftBase((20, 30), trj.golden_angle_2d(30, 48), Jd=6, Kd=(48, 64), Ld=1024, device=dev())
    assert autotune.tune_fft(C) == ("auto", {})  # no fused passes for K = 48: nothing to choose
