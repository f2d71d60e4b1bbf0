"""mi355nufft_table1_adjoint_noalloc: the raw table1 scatter with its bin
sort in the library's static device scratch (no allocation in the call).

Checked through the C ABI against mi355nufft_table1_adjoint (the entry
point it is a variant of) on two calls in a row, the second with fewer
points at another address, in complex64 and complex128, 2-D and 3-D; an
input beyond the scratch is refused with a clear error before anything is
enqueued.
"""

import ctypes

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import _lib
from mrrt.nufft_amd import trajectories as trj
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {np.float32: 1e-5, np.float64: 1e-12}


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _fn():
    lib = _lib.get_lib()
    return lib, lib.mi355nufft_table1_adjoint_noalloc


def _tables(Nd, Jd, Kd, L, rdt):
    hs = [np.real(O.table_make1("fast", Nd[d], Jd[d], Kd[d], L, "real")[0]).astype(rdt) for d in range(len(Kd))]
    hd = [torch.from_numpy(h).to(dev()) for h in hs]
    return hs, hd, (ctypes.c_void_p * len(Kd))(*[h.data_ptr() for h in hd])


def _call(fn, Kd, Jd, L, rdt, fm_t, hp, tm_t, nreps):
    M = tm_t.shape[1]
    out = torch.empty((nreps, int(np.prod(Kd))), dtype=fm_t.dtype, device=dev())
    prec = 0 if rdt == np.float32 else 1
    rc = fn(len(Kd), _lib.i64arr(Kd), _lib.i32arr(Jd), L, prec, 0, _vp(fm_t), hp, _vp(tm_t), M, _vp(out), nreps,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return rc, out


def _tm(om, Kd, rdt):
    om = om.astype(rdt)
    return np.stack([om[:, d] / rdt(2 * np.pi / Kd[d]) for d in range(len(Kd))]).astype(rdt)


@pytest.mark.parametrize("rdt", [np.float32, np.float64])
@pytest.mark.parametrize("ndim", [2, 3])
def test_noalloc_matches_adjoint_two_calls(rdt, ndim):
    lib, fn = _fn()
    if ndim == 2:
        Nd, Kd, Jd, L = (256, 256), (512, 512), (6, 6), 1024
        om1 = trj.golden_angle_2d(402, 512)
        om2 = trj.radial_2d(60, 300) * 0.8
    else:
        Nd, Kd, Jd, L = (48, 48, 48), (96, 96, 96), (6, 6, 6), 512
        om1 = trj.kooshball_3d(1500, 96, seed=1)
        om2 = trj.kooshball_3d(300, 64, seed=2) * 0.9
    hs, hd, hp = _tables(Nd, Jd, Kd, L, rdt)
    cdt = np.complex64 if rdt == np.float32 else np.complex128
    rng = np.random.default_rng(ndim)
    nreps = 2
    for om in (om1, om2):   # second call: fewer points, another address
        tm = _tm(om, Kd, rdt)
        tm_t = torch.from_numpy(tm).to(dev())
        fm = (rng.standard_normal((nreps, tm.shape[1])) + 1j * rng.standard_normal((nreps, tm.shape[1]))).astype(cdt)
        fm_t = torch.from_numpy(fm).to(dev())
        rc, got = _call(fn, Kd, Jd, L, rdt, fm_t, hp, tm_t, nreps)
        _lib.check(rc)
        ref = torch.empty_like(got)
        prec = 0 if rdt == np.float32 else 1
        _lib.check(lib.mi355nufft_table1_adjoint(len(Kd), _lib.i64arr(Kd), _lib.i32arr(Jd), L, prec, 0, _vp(fm_t),
                                                 hp, _vp(tm_t), tm.shape[1], _vp(ref), nreps,
                                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        torch.cuda.synchronize()
        g, r = got.cpu().numpy(), ref.cpu().numpy()
        assert rel_l2(g, r) <= TOL[rdt]
        if ndim == 2 and om is om2:
            assert rel_l2(g, O.interp_adjoint(fm.T, Kd, hs, Jd, L, tm.T).T) <= TOL[rdt]


def test_noalloc_same_points_twice_and_refusal():
    """Unchanged points reuse the scratch's sort; an input beyond the scratch
    is refused (ValueError from EINVAL) and leaves the output untouched."""
    lib, fn = _fn()
    Nd, Kd, Jd, L = (64, 64), (128, 128), (6, 6), 1024
    hs, hd, hp = _tables(Nd, Jd, Kd, L, np.float32)
    tm = _tm(trj.golden_angle_2d(64, 128), Kd, np.float32)
    tm_t = torch.from_numpy(tm).to(dev())
    rng = np.random.default_rng(5)
    fm = (rng.standard_normal((1, tm.shape[1])) + 1j * rng.standard_normal((1, tm.shape[1]))).astype(np.complex64)
    fm_t = torch.from_numpy(fm).to(dev())
    ref = O.interp_adjoint(fm.T, Kd, hs, Jd, L, tm.T).T
    for _ in range(2):
        rc, got = _call(fn, Kd, Jd, L, np.float32, fm_t, hp, tm_t, 1)
        _lib.check(rc)
        torch.cuda.synchronize()
        assert rel_l2(got.cpu().numpy(), ref) <= 1e-5
    # 2-D float: 28 bytes of scratch per point, 64 MiB hold fewer than 2.4 M points
    M_big = 3_000_000
    tmb = torch.zeros((2, M_big), dtype=torch.float32, device=dev())
    fmb = torch.zeros((1, M_big), dtype=torch.complex64, device=dev())
    out = torch.full((1, int(np.prod(Kd))), 7.0 + 0j, dtype=torch.complex64, device=dev())
    rc = fn(2, _lib.i64arr(Kd), _lib.i32arr(Jd), L, 0, 0, _vp(fmb), hp, _vp(tmb), M_big, _vp(out), 1,
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 1
    msg = lib.mi355nufft_last_error().decode()
    assert "scratch" in msg and "mi355nufft_table1_adjoint" in msg
    torch.cuda.synchronize()
    assert torch.all(out == (7.0 + 0j)).item()
