"""Raw table1 entry points (the reference's interp_table / interp_table_adj
plug-in boundary, mrrt/nufft/_nufft.py:343-344): unsorted points, run on the
tiled kernels through a bin sort that is cached per (device, stream) and
re-made when tm changes, also in place at the same address.

Checked against the oracle's table1 kernels (nufft_table.template.c
restated) and, for the real-kernel gather, bit for bit against the
reference's own C kernels (oracle/_ref) at the C2 size.
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


class Raw:
    """Device copies of one geometry's tables and the raw calls."""

    def __init__(self, Nd, Kd, Jd, L, rdt, phasing="real"):
        self.Nd, self.Kd, self.Jd, self.L, self.rdt = Nd, Kd, Jd, L, rdt
        self.ndim = len(Kd)
        self.cdt = np.complex64 if rdt == np.float32 else np.complex128
        self.ck = phasing == "complex"
        self.hs = []
        for d in range(self.ndim):
            h, _ = O.table_make1("fast", Nd[d], Jd[d], Kd[d], L, phasing)
            self.hs.append(h.astype(self.cdt) if self.ck else np.real(h).astype(rdt))
        self.hd = [torch.from_numpy(h).to(dev()) for h in self.hs]
        self.hp = (ctypes.c_void_p * self.ndim)(*[h.data_ptr() for h in self.hd])
        self.lib = _lib.get_lib()
        self.prec = 0 if rdt == np.float32 else 1

    def _geom(self):
        return (self.ndim, _lib.i64arr(self.Kd), _lib.i32arr(self.Jd), self.L, self.prec, int(self.ck))

    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def forward(self, ck_t, tm_t, nreps):
        M = tm_t.shape[1]
        fm = torch.empty((nreps, M), dtype=ck_t.dtype, device=dev())
        _lib.check(self.lib.mi355nufft_table1_forward(*self._geom(), _vp(ck_t), self.hp, _vp(tm_t), M, _vp(fm),
                                                      nreps, self.stream()))
        return fm

    def adjoint(self, fm_t, tm_t, nreps):
        M = tm_t.shape[1]
        out = torch.empty((nreps, int(np.prod(self.Kd))), dtype=fm_t.dtype, device=dev())
        _lib.check(self.lib.mi355nufft_table1_adjoint(*self._geom(), _vp(fm_t), self.hp, _vp(tm_t), M, _vp(out),
                                                      nreps, self.stream()))
        return out

    def oracle_fwd(self, ck, tm):
        return O.interp_forward(ck.T, self.Kd, self.hs, self.Jd, self.L, tm.T, self.ck).T

    def oracle_adj(self, fm, tm):
        return O.interp_adjoint(fm.T, self.Kd, self.hs, self.Jd, self.L, tm.T, self.ck).T


def _tm(om, Kd, rdt):
    """(ndim, M) tm = omega / (2 pi / K) in the real dtype (_nufft.py:338-342)."""
    om = om.astype(rdt)
    return np.stack([om[:, d] / rdt(2 * np.pi / Kd[d]) for d in range(len(Kd))]).astype(rdt)


@pytest.mark.parametrize("rdt", [np.float32, np.float64])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_raw_cached_sort_follows_in_place_updates(rdt, phasing):
    """Repeated calls on the same tm give the same result; rewriting tm in
    place (same device address) is seen by the next call; a different tm of
    another size at another address too."""
    Nd, Kd, Jd, L = (40, 36), (80, 72), (6, 6), 1024
    R = Raw(Nd, Kd, Jd, L, rdt, phasing)
    om = trj.golden_angle_2d(40, 80)
    tm = _tm(om, Kd, rdt)
    tm_t = torch.from_numpy(tm).to(dev())
    rng = np.random.default_rng(1)
    nreps = 3
    ck = (rng.standard_normal((nreps, int(np.prod(Kd)))) + 1j * rng.standard_normal((nreps, int(np.prod(Kd))))
          ).astype(R.cdt)
    ck_t = torch.from_numpy(ck).to(dev())
    fm = (rng.standard_normal((nreps, tm.shape[1])) + 1j * rng.standard_normal((nreps, tm.shape[1]))).astype(R.cdt)
    fm_t = torch.from_numpy(fm).to(dev())
    y1 = R.forward(ck_t, tm_t, nreps).cpu().numpy()
    y1b = R.forward(ck_t, tm_t, nreps).cpu().numpy()
    np.testing.assert_array_equal(y1, y1b)
    assert rel_l2(y1, R.oracle_fwd(ck, tm)) <= TOL[rdt]
    x1 = R.adjoint(fm_t, tm_t, nreps).cpu().numpy()
    assert rel_l2(x1, R.oracle_adj(fm, tm)) <= TOL[rdt]
    # in place: same address, new values
    tm2 = _tm(om[::-1] * 0.7, Kd, rdt)
    tm_t.copy_(torch.from_numpy(np.ascontiguousarray(tm2)).to(dev()))
    y2 = R.forward(ck_t, tm_t, nreps).cpu().numpy()
    assert rel_l2(y2, R.oracle_fwd(ck, tm2)) <= TOL[rdt]
    x2 = R.adjoint(fm_t, tm_t, nreps).cpu().numpy()
    assert rel_l2(x2, R.oracle_adj(fm, tm2)) <= TOL[rdt]
    # a single changed coordinate is enough
    tm3 = tm2.copy()
    tm3[1, 17] = tm3[1, 17] + rdt(3.25)
    tm_t.copy_(torch.from_numpy(tm3).to(dev()))
    y3 = R.forward(ck_t, tm_t, nreps).cpu().numpy()
    assert rel_l2(y3, R.oracle_fwd(ck, tm3)) <= TOL[rdt]
    assert np.abs(y3[:, 17] - y2[:, 17]).max() > 0
    # another point set (smaller, elsewhere)
    tm4 = _tm(trj.radial_2d(8, 40), Kd, rdt)
    tm4_t = torch.from_numpy(tm4).to(dev())
    y4 = R.forward(ck_t, tm4_t, nreps).cpu().numpy()
    assert rel_l2(y4, R.oracle_fwd(ck, tm4)) <= TOL[rdt]


@pytest.mark.parametrize("rdt", [np.float32, np.float64])
def test_raw_c2_points_vs_reference_kernels(rdt):
    """The C2 golden-angle points (205,824, Kd 512^2) through the raw entry
    points: the real-kernel gather equals the reference's own C kernel
    (oracle/_ref, built from nufft_table.c) bit for bit in float32, and in
    float64 up to the last bits of a handful of samples; the scatter equals
    the oracle to the parity bar."""
    from oracle import ref_ctypes
    Nd, Kd = (256, 256), (512, 512)
    om = trj.golden_angle_2d(402, 512)
    R = Raw(Nd, Kd, (6, 6), 1024, rdt)
    tm = _tm(om, Kd, rdt)
    tm_t = torch.from_numpy(tm).to(dev())
    rng = np.random.default_rng(2)
    nk = int(np.prod(Kd))
    ck = (rng.standard_normal((1, nk)) + 1j * rng.standard_normal((1, nk))).astype(R.cdt)
    y_raw = R.forward(torch.from_numpy(ck).to(dev()), tm_t, 1).cpu().numpy()
    assert rel_l2(y_raw, R.oracle_fwd(ck, tm)) <= TOL[rdt]
    if ref_ctypes.available():
        y_ref = ref_ctypes.ref_interp_forward(ck.T, Kd, R.hs, (6, 6), 1024, tm.T).T
        if rdt == np.float32:
            np.testing.assert_array_equal(y_raw, y_ref)
        else:
            # the reference is built with -ffast-math (setup.py:110), which
            # lets gcc reorder double arithmetic: a handful of samples differ
            # in the last bits (1.9e-16 absolute on values of order 1)
            diff = np.abs(y_raw - y_ref)
            assert diff.max() <= 4 * np.finfo(np.float64).eps * np.abs(y_ref).max()
            assert np.count_nonzero(diff) <= 1e-4 * y_ref.size
    fm = (rng.standard_normal((1, tm.shape[1])) + 1j * rng.standard_normal((1, tm.shape[1]))).astype(R.cdt)
    x_raw = R.adjoint(torch.from_numpy(fm).to(dev()), tm_t, 1).cpu().numpy()
    assert rel_l2(x_raw, R.oracle_adj(fm, tm)) <= TOL[rdt]


def test_raw_two_streams():
    """Calls on two streams use separate cached sorts and both are right."""
    Nd, Kd, Jd, L = (32, 32), (64, 64), (6, 6), 1024
    R = Raw(Nd, Kd, Jd, L, np.float32)
    tmA = _tm(trj.golden_angle_2d(24, 64), Kd, np.float32)
    tmB = _tm(trj.radial_2d(16, 64) * 0.9, Kd, np.float32)
    rng = np.random.default_rng(3)
    nk = int(np.prod(Kd))
    ck = (rng.standard_normal((1, nk)) + 1j * rng.standard_normal((1, nk))).astype(np.complex64)
    ck_t = torch.from_numpy(ck).to(dev())
    tA, tB = torch.from_numpy(tmA).to(dev()), torch.from_numpy(tmB).to(dev())
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):
            yA = R.forward(ck_t, tA, 1)
        with torch.cuda.stream(s2):
            yB = R.forward(ck_t, tB, 1)
        torch.cuda.synchronize()
        outs.append((yA.cpu().numpy(), yB.cpu().numpy()))
    for yA, yB in outs:
        assert rel_l2(yA, R.oracle_fwd(ck, tmA)) <= 1e-5
        assert rel_l2(yB, R.oracle_fwd(ck, tmB)) <= 1e-5
