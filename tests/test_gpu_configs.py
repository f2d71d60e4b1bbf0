"""GPU parity on the five benchmark configurations, at their benchmark sizes.

Each test builds its configuration with ``mrrt.nufft_amd.workloads`` (the
inputs the benchmark times) and compares the HIP path with the oracle, the
restatement of the reference's nufft_forward / nufft_adj
(mrrt/nufft/_nufft.py:1275-1397, 1459-1578), and with the exact FP64 NUDFT
(mrrt/nufft/_dtft.py) on seeded subsets.  Where the oracle would take
minutes (3-D), the comparison runs on a seeded subset of samples and the
full size is covered by size-independent properties (adjointness, graph
replay == eager).

Bars: relative L2 <= 1e-5 (complex64) and <= 1e-12 (complex128) against the
oracle; against the exact NUDFT the reference's own tolerances
(mrrt/nufft/tests/test_nufft.py: rtol 1e-3 in 2-D, 1e-2 in 3-D).  Every
measured error is printed as ``CONFIG-ERROR <config> <quantity> <value>``.
"""

import numpy as np
import pytest
import torch

from mrrt.nufft_amd import NufftBase
from mrrt.nufft_amd import workloads as W
from mrrt.nufft_amd.autograd import nufft
from mrrt.nufft_amd.batched import NufftFrames
from mrrt.nufft_amd.mri import SenseOperator
from oracle import nufft_oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"single": 1e-5, "double": 1e-12}


def dev():
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    return torch.device("cuda", 0)


def report(config, what, value):
    print("CONFIG-ERROR %s %s %.3e" % (config, what, value), flush=True)
    return value


def _ops(wl, **kw):
    geo = dict(Jd=wl["Jd"], Kd=wl["Kd"], Ld=wl["Ld"], precision=wl["precision"])
    geo.update(kw)
    return geo


# ----------------------------------------------------------------- C1 / C2
@pytest.mark.parametrize("config,precision", [("c1", "single"), ("c2", "single"), ("c2", "double")])
def test_2d_config_full_size_vs_oracle_and_nudft(config, precision):
    """C1 (Nd 128^2, 64 x 256 radial) and C2 (Nd 256^2, Kd 512^2, 402 x 512
    golden-angle, complex64 and complex128): forward and adjoint at full
    size against the oracle; forward on 512 seeded samples and adjoint on 512
    seeded pixels against the FP64 NUDFT."""
    wl = W.c1() if config == "c1" else W.c2(precision)
    tag = "%s-%s" % (config, "c64" if precision == "single" else "c128")
    geo = _ops(wl, precision=precision)
    A = NufftBase(wl["Nd"], wl["om"], device=dev(), **geo)
    B = O.NufftOracle(wl["Nd"], wl["om"], **geo)
    x = wl["x"]
    y = A.fft(x)
    yo = B.fft(x)
    e_f = report(tag, "forward-vs-oracle", rel_l2(y, yo))
    rng = np.random.default_rng(11)
    ya = (rng.standard_normal(yo.shape) + 1j * rng.standard_normal(yo.shape)).astype(B.cdt)
    xa = A.adj(ya)
    xo = B.adj(ya)
    e_a = report(tag, "adjoint-vs-oracle", rel_l2(xa, xo))
    assert e_f <= TOL[precision] and e_a <= TOL[precision]
    sel = rng.choice(wl["om"].shape[0], 512, replace=False)
    ye = O.dtft(x.ravel(order="F"), wl["om"][sel], wl["Nd"], np.zeros(2))
    e_fn = report(tag, "forward-vs-nudft", rel_l2(y[sel], ye))
    pix = rng.choice(int(np.prod(wl["Nd"])), 512, replace=False)
    xe = O.dtft_adj(ya, wl["om"], wl["Nd"], np.zeros(2), pixels=pix)
    e_an = report(tag, "adjoint-vs-nudft", rel_l2(xa.ravel(order="F")[pix], xe))
    assert e_fn < 1e-3 and e_an < 1e-3


def test_c2_full_size_adjoint_fixed_point_vs_float64():
    """C2 adjoint of 205,824 random samples: the default 32-bit fixed-point
    scatter tile against the float64 tile, and both against the oracle."""
    wl = W.c2("single")
    geo = _ops(wl, precision="single")
    A = NufftBase(wl["Nd"], wl["om"], device=dev(), **geo)
    B = O.NufftOracle(wl["Nd"], wl["om"], **geo)
    rng = np.random.default_rng(16)
    y = (rng.standard_normal(A.M) + 1j * rng.standard_normal(A.M)).astype(np.complex64)
    assert A.plan_info()["fixed_point_scatter"]
    xf = A.adj(y)
    A.set_scatter("float64")
    xd = A.adj(y)
    xo = B.adj(y)
    e = report("c2-c64", "adjoint-fixed-point-vs-float64-tile(full size)", rel_l2(xf, xd))
    e_f = report("c2-c64", "adjoint-fixed-point-vs-oracle", rel_l2(xf, xo))
    e_d = report("c2-c64", "adjoint-float64-tile-vs-oracle", rel_l2(xd, xo))
    assert e <= 5e-6 and e_f <= 1e-5 and e_d <= 1e-5


# ---------------------------------------------------------------------- C3
@pytest.fixture(scope="module")
def c3_setup():
    wl = W.c3()
    A = NufftBase(wl["Nd"], wl["om"], device=dev(), **_ops(wl))
    return wl, A


def test_c3_forward_on_sample_subset(c3_setup):
    """C3 (Nd 128^3, Kd 256^3, kooshball 8192 x 256): the full transform on
    the GPU; 4,096 seeded samples against the oracle (full 256^3 grid FFT),
    64 of them against the FP64 NUDFT."""
    wl, A = c3_setup
    sel = np.sort(np.random.default_rng(12).choice(A.M, 4096, replace=False))
    y = A.fft(wl["x"])
    B = O.NufftOracle(wl["Nd"], wl["om"][sel], **_ops(wl))
    yo = B.fft(wl["x"])
    e = report("c3-c64", "forward-vs-oracle(4096 samples)", rel_l2(y[sel], yo))
    assert e <= TOL["single"]
    ye = O.dtft(wl["x"].ravel(order="F"), wl["om"][sel[:64]], wl["Nd"], np.zeros(3))
    en = report("c3-c64", "forward-vs-nudft(64 samples)", rel_l2(y[sel[:64]], ye))
    assert en < 1e-2


def test_c3_adjoint_on_sample_subset(c3_setup):
    """C3 adjoint over all 2,097,152 points with y zero outside 4,096 seeded
    samples, against the oracle's adjoint of those samples; the exact NUDFT
    on 256 seeded pixels."""
    wl, A = c3_setup
    rng = np.random.default_rng(13)
    sel = np.sort(rng.choice(A.M, 4096, replace=False))
    ys = (rng.standard_normal(4096) + 1j * rng.standard_normal(4096)).astype(np.complex64)
    y = np.zeros(A.M, np.complex64)
    y[sel] = ys
    xa = A.adj(y)
    B = O.NufftOracle(wl["Nd"], wl["om"][sel], **_ops(wl))
    xo = B.adj(ys)
    e = report("c3-c64", "adjoint-vs-oracle(4096 samples)", rel_l2(xa, xo))
    assert e <= TOL["single"]
    pix = rng.choice(int(np.prod(wl["Nd"])), 256, replace=False)
    xe = O.dtft_adj(ys, wl["om"][sel], wl["Nd"], np.zeros(3), pixels=pix)
    en = report("c3-c64", "adjoint-vs-nudft(256 pixels)", rel_l2(xa.ravel(order="F")[pix], xe))
    assert en < 1e-2


def test_c3_full_size_adjointness(c3_setup):
    """<A x, y> == <x, A^H y> over the full C3 operator (float64 sums of the
    complex64 results)."""
    wl, A = c3_setup
    rng = np.random.default_rng(14)
    x = wl["x"].astype(np.complex64)
    y = (rng.standard_normal(A.M) + 1j * rng.standard_normal(A.M)).astype(np.complex64)
    lhs = np.vdot(A.fft(x).astype(np.complex128), y.astype(np.complex128))
    rhs = np.vdot(x.astype(np.complex128), A.adj(y).astype(np.complex128))
    e = report("c3-c64", "adjointness", abs(lhs - rhs) / abs(lhs))
    assert e <= 1e-5


def test_c3_full_size_adjoint_fixed_point_vs_float64(c3_setup):
    """C3 adjoint of 2,097,152 random samples: the default 32-bit fixed-point
    scatter tile against the float64 tile (the one the subset test pins to
    the oracle).  This is the fixed-point tile's error at full load."""
    wl, A = c3_setup
    rng = np.random.default_rng(15)
    y = (rng.standard_normal(A.M) + 1j * rng.standard_normal(A.M)).astype(np.complex64)
    assert A.plan_info()["fixed_point_scatter"]
    xf = A.adj(y)
    A.set_scatter("float64")
    try:
        xd = A.adj(y)
    finally:
        A.set_scatter("auto")
    e = report("c3-c64", "adjoint-fixed-point-vs-float64-tile(full size)", rel_l2(xf, xd))
    assert e <= 1e-5


# ---------------------------------------------------------------------- C4
@pytest.fixture(scope="module")
def c4_setup():
    wl = W.c4()
    F, C = len(wl["frames"]), wl["ncoils"]
    Fr = NufftFrames(wl["Nd"], wl["frames"], C, Jd=6, Kd=wl["Kd"], Ld=wl["Ld"], device=dev())
    return wl, Fr, F, C


def test_c4_frames_vs_oracle(c4_setup):
    """C4 (24 frames x 32 coils, 13..89 spokes x 512 per frame): the whole
    batch in one plan; 3 seeded frames x 2 seeded coils against the oracle,
    forward and adjoint."""
    wl, Fr, F, C = c4_setup
    g = torch.Generator(device=dev()).manual_seed(0)
    x = torch.randn((F, C) + wl["Nd"], dtype=torch.complex64, device=dev(), generator=g)
    y = Fr.forward(x)
    yt = torch.randn((C, Fr.M), dtype=torch.complex64, device=dev(), generator=g)
    xa = Fr.adjoint(yt)
    rng = np.random.default_rng(15)
    worst_f = worst_a = 0.0
    for f in rng.choice(F, 3, replace=False):
        R = O.NufftOracle(wl["Nd"], wl["frames"][f], **_ops(wl))
        sl = Fr.frame_slice(f)
        for c in rng.choice(C, 2, replace=False):
            worst_f = max(worst_f, rel_l2(y[c, sl].cpu().numpy(), R.fft(x[f, c].cpu().numpy())))
            worst_a = max(worst_a, rel_l2(xa[f, c].cpu().numpy(), R.adj(yt[c, sl].cpu().numpy())))
    report("c4-c64", "forward-vs-oracle(worst of 3 frames x 2 coils)", worst_f)
    report("c4-c64", "adjoint-vs-oracle(worst of 3 frames x 2 coils)", worst_a)
    assert worst_f <= TOL["single"] and worst_a <= TOL["single"]


def test_c4_graph_replays_match_eager_full_size(c4_setup):
    """The captured forward and adjoint of all 768 images, replayed three
    times with eager work and synchronisations between the replays, equal
    the eager transforms; afterwards an eager adjoint of y = 0 gives exact
    zeros (the captured adjoint left the plan's grid clear)."""
    wl, Fr, F, C = c4_setup
    g = torch.Generator(device=dev()).manual_seed(1)
    x = torch.randn((F, C) + wl["Nd"], dtype=torch.complex64, device=dev(), generator=g)
    y_e = Fr.forward(x).clone()
    xa_e = Fr.adjoint(y_e).clone()
    y = torch.empty_like(y_e)
    xa = torch.empty_like(x)
    Fr.capture(x, y, xa)
    worst = 0.0
    for rep in range(3):
        y.zero_()
        xa.zero_()
        Fr.replay_forward()
        Fr.replay_adjoint()
        torch.cuda.synchronize()
        worst = max(worst, rel_l2(y.cpu().numpy(), y_e.cpu().numpy()),
                    rel_l2(xa.cpu().numpy(), xa_e.cpu().numpy()))
        _ = Fr.forward(x)   # eager work between the replays
        torch.cuda.synchronize()
    report("c4-c64", "graph-replay-vs-eager", worst)
    assert worst <= 1e-6
    z = Fr.adjoint(torch.zeros_like(y_e))
    torch.cuda.synchronize()
    assert torch.count_nonzero(z).item() == 0


def test_c4_sense_frames_vs_oracle(c4_setup):
    """SENSE over the frames plan with the configuration's complex coil maps:
    E x (one image per frame) and E^H y against per-frame oracle loops on 2
    seeded frames."""
    wl, Fr, F, C = c4_setup
    E = SenseOperator(Fr, wl["smaps"])
    x = torch.from_numpy(wl["x"].astype(np.complex64)).to(dev())
    y = E.forward_dev(x)
    rng = np.random.default_rng(16)
    yt = (rng.standard_normal((C, Fr.M)) + 1j * rng.standard_normal((C, Fr.M))).astype(np.complex64)
    xa = E.adjoint_dev(torch.from_numpy(yt).to(dev())).cpu().numpy()
    yn = y.cpu().numpy()
    smaps = wl["smaps"].astype(np.complex64)
    worst_f = worst_a = 0.0
    for f in rng.choice(F, 2, replace=False):
        R = O.NufftOracle(wl["Nd"], wl["frames"][f], **_ops(wl))
        sl = Fr.frame_slice(f)
        ref_y = np.stack([R.fft(smaps[c] * wl["x"][f].astype(np.complex64)) for c in range(C)])
        worst_f = max(worst_f, rel_l2(yn[:, sl], ref_y))
        ref_x = sum(np.conj(smaps[c]) * R.adj(yt[c, sl]) for c in range(C))
        worst_a = max(worst_a, rel_l2(xa[f], ref_x))
    report("c4-c64", "sense-forward-vs-oracle(worst of 2 frames)", worst_f)
    report("c4-c64", "sense-adjoint-vs-oracle(worst of 2 frames)", worst_a)
    assert worst_f <= TOL["single"] and worst_a <= TOL["single"]


# ---------------------------------------------------------------------- C5
def _crossing(omega, Kd, J, L, h):
    """True where omega +- h moves some tap into another table cell."""
    mask = np.zeros(omega.shape, bool)
    for d in range(omega.shape[1]):
        cells = []
        for sgn in (-1, 1):
            t = (omega[:, d] + sgn * h) / (2 * np.pi / Kd[d])
            k0 = 1 + np.floor(t - J / 2.0)
            p = (t[:, None] - (k0[:, None] + np.arange(J))) * L
            cells.append(np.floor(p) + 1e6 * k0[:, None])
        mask[:, d] = np.any(cells[0] != cells[1], axis=1)
    return mask


def test_c5_autograd_full_size():
    """C5 (16 x 4096 spiral, 8 coils, loss ||A(x, om) - y||^2, y from the
    perturbed trajectory), complex64 as configured:
      d/dx  == 2 A^H (A x - y) with the oracle's adjoint (1e-5);
      d/dom of the complex128 operator == FP64 central finite differences
            (h = 1e-6) on 24 seeded entries;
      d/dom of complex64 vs complex128 (relative L2, bounded by the
            condition of the residual as in the small-size test)."""
    wl = W.c5()
    C, Nd, om = wl["ncoils"], wl["Nd"], wl["om"]
    H = 1e-6
    op32 = NufftBase(Nd, om, device=dev(), **_ops(wl))
    op64 = NufftBase(Nd, om, device=dev(), **_ops(wl, precision="double"))
    with torch.no_grad():
        y_t = nufft(op64, torch.from_numpy(wl["x"]).to(dev()), torch.from_numpy(wl["om_pert"]).to(dev()))
    # complex64 gradient with respect to x vs the oracle adjoint of the residual
    x32 = torch.from_numpy(wl["x"].astype(np.complex64)).to(dev()).requires_grad_(True)
    w32 = torch.from_numpy(om.astype(np.float32)).to(dev()).requires_grad_(True)
    y32 = nufft(op32, x32, w32)
    L32 = ((y32 - y_t.to(torch.complex64)).abs() ** 2).sum()
    gx32, gw32 = torch.autograd.grad(L32, (x32, w32))
    r = (y32.detach() - y_t.to(torch.complex64)).cpu().numpy()
    # the points were set from the float32 omega, so phase_after is the
    # reference's float32 one (_nufft.py:313, :717-724): the oracle gets the
    # same float32 omega
    B = O.NufftOracle(Nd, om.astype(np.float32), **_ops(wl))
    ref_gx = np.stack([2 * B.adj(r[c]) for c in range(C)])
    e_x = report("c5-c64", "grad-x-vs-oracle-adjoint", rel_l2(gx32.cpu().numpy(), ref_gx))
    assert e_x <= TOL["single"]
    # complex128 trajectory gradient vs FP64 central finite differences
    x64 = torch.from_numpy(wl["x"]).to(dev())
    om64 = torch.from_numpy(om).to(dev())
    wg = om64.clone().requires_grad_(True)

    def loss64(w):
        return ((nufft(op64, x64, w) - y_t).abs() ** 2).sum()

    (gw64,) = torch.autograd.grad(loss64(wg), (wg,))
    gwn = gw64.cpu().numpy()
    scale = np.abs(gwn).max()
    cross = _crossing(om, wl["Kd"], 6, wl["Ld"], H)
    worst, checked = 0.0, 0
    for m in np.random.default_rng(17).choice(om.shape[0], 16, replace=False):
        for d in range(2):
            if cross[m, d]:
                continue
            e = torch.zeros_like(om64)
            e[m, d] = H
            with torch.no_grad():
                fd = (loss64(om64 + e) - loss64(om64 - e)).item() / (2 * H)
            err = abs(fd - gwn[m, d]) / max(abs(gwn[m, d]), 1e-3 * scale)
            worst = max(worst, err)
            checked += 1
    report("c5-c128", "grad-omega-vs-fp64-fd(worst relative)", worst)
    assert checked >= 16 and worst <= 1e-5
    # complex64 vs the FD-checked complex128 gradients
    with torch.no_grad():
        y64 = nufft(op64, x64, om64)
        kappa = (torch.linalg.vector_norm(y64) / torch.linalg.vector_norm(y64 - y_t)).item()
    e_w = report("c5-c64", "grad-omega-vs-c128", rel_l2(gw32.double().cpu().numpy(), gwn))
    report("c5-c64", "residual-condition-kappa", kappa)
    assert 1 < kappa < 1e4
    assert e_w <= 2e-4 * kappa
