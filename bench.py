"""Benchmark: table1 NUFFT forward+adjoint pairs/s on one MI355X.

Workload (one "step" = one forward + one adjoint transform, end to end
including rocFFT, inputs resident in HBM):
    2-D, Nd = (256, 256), Kd = (512, 512), Jd = (6, 6), Ld = 1024,
    golden-angle radial 402 spokes x 512 readout (M = 205,824), complex64,
    image = complex Gaussian + 10 random ellipses (seeded).

Besides the JSON line's headline value, the line carries
  roofline      the dominant interpolation kernel's achieved algorithmic
                HBM bandwidth (bytes per launch / mean launch time, HIP
                events on the plan's stream) against 8 TB/s,
  cpu_baseline  the reference's own C kernels (oracle/_ref, built from the
                reference sources) with numpy FFTs, on a bounded sample,
  stages        per-stage mean times, and the complex128 and 3-D runs.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0

KERNEL_NAMES = {"interp": "interp_tiled_kernel<float, 2, false, 6, true, false, true>",
                "spread": "spread_tiled_kernel<float, 2, false, 6, true, true, false, 1>"}

# HBM bytes per launch of the 2-D 256^2 complex64 kernels, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over this bench (--calibrate-copy), each
# scaled by what the counter reads for a 256 MiB device copy (FETCH_SIZE
# x 2.0, WRITE_SIZE x 1.0); profiles/nufft2d_256_c64_hbm_traffic.md.
TRAFFIC_BYTES = {"spread": 18007664, "interp": 12857829}
# The same for the 3-D 128^3 / 256^3 kooshball kernels (16^3 bins, fixed-point
# scatter tile; profiles/nufft3d_128_c64_hbm_traffic.md) and the 24-frame x
# 32-coil batch (profiles/nufft2d_256_c64_frames_kernel_stats_and_traffic.md).
TRAFFIC_BYTES_C3 = {"interp": 369.2e6, "spread": 516.4e6}
TRAFFIC_BYTES_C4 = {"interp": 2468.2e6, "spread": 2102.3e6}  # 64x32 bins, 1,024 points per subproblem


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the complex128 / 3-D extra measurements")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--graph", action="store_true", help="time a HIP-graph replay instead of eager launches")
    ap.add_argument("--fft", choices=("tuned", "fused", "rocfft"), default="tuned",
                    help="grid FFT: the faster of the two measured here (default), the fused pruned passes, or rocFFT")
    ap.add_argument("--cpu-only", action="store_true",
                    help="only the CPU baseline: the reference pipeline (its C kernels + scipy.fft) per stage on the "
                         "configurations of --cpu-configs, OMP_NUM_THREADS threads; one JSON line per configuration "
                         "(no GPU needed)")
    ap.add_argument("--cpu-configs", default="c1,c2,c2_c128,c3,c4,c5")
    ap.add_argument("--cpu-runs", type=int, default=10, help="--cpu-only: timed runs per configuration (median)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run three 256 MiB device-to-device copies (a known-size reference for "
                         "FETCH_SIZE / WRITE_SIZE counter passes)")
    return ap.parse_args()


def c2_workload(precision="single"):
    """The 2-D benchmark configuration (mrrt.nufft_amd.workloads.c2)."""
    from mrrt.nufft_amd import workloads as W
    return W.c2(precision)


def c3_workload():
    """The 3-D benchmark configuration (mrrt.nufft_amd.workloads.c3)."""
    from mrrt.nufft_amd import workloads as W
    return W.c3()


def interp_bytes(M, ndim, nk, nreps, real_size):
    """Algorithmic HBM bytes of one table1 gather (or scatter) launch:
    tm read once, samples written (read), grid read (written) once."""
    return M * ndim * real_size + nreps * M * 2 * real_size + nreps * nk * 2 * real_size


class GpuRun:
    def __init__(self, wl, mode="table"):
        import torch
        from mrrt.nufft_amd import NufftBase
        self.torch = torch
        self.dev = torch.device("cuda", 0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        self.A = NufftBase(wl["Nd"], wl["om"], Jd=wl["Jd"], Kd=wl["Kd"], Ld=wl["Ld"], precision=wl["precision"],
                           mode=mode, device=self.dev)
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0
        cdt = torch.complex64 if wl["precision"] == "single" else torch.complex128
        nd = len(wl["Nd"])
        xr = torch.from_numpy(np.ascontiguousarray(wl["x"])).to(self.dev).to(cdt)
        # Fortran-ordered image Nd + (1,), resident on the device
        self.x = xr.permute(*reversed(range(nd))).contiguous().permute(*reversed(range(nd))).unsqueeze(-1)
        self.M = self.A.M
        self.nk = int(np.prod(wl["Kd"]))
        self.ndim = nd
        self.real_size = 4 if wl["precision"] == "single" else 8
        self.y = torch.empty((1, self.M), dtype=cdt, device=self.dev)
        self.grid = torch.empty((1, self.nk), dtype=cdt, device=self.dev)
        self.xo = torch.empty((1,) + tuple(reversed(wl["Nd"])), dtype=cdt, device=self.dev)
        self.xo = self.xo.permute(*(tuple(range(nd, 0, -1)) + (0,)))
        import ctypes
        from mrrt.nufft_amd import _lib
        self.ct = ctypes
        self.L = _lib
        self.lib = _lib.get_lib()
        self.xs = _lib.i64arr(list(self.x.stride()))
        self.xos = _lib.i64arr(list(self.xo.stride()))
        self.p = self.A._plan
        vp = ctypes.c_void_p
        self.px, self.py, self.pg, self.pxo = (vp(self.x.data_ptr()), vp(self.y.data_ptr()),
                                               vp(self.grid.data_ptr()), vp(self.xo.data_ptr()))

    def pair(self):
        chk, lib, p = self.L.check, self.lib, self.p
        chk(lib.mi355nufft_forward(p, self.px, self.xs, self.py, 1, self.pg))
        chk(lib.mi355nufft_adjoint(p, self.py, self.pxo, self.xos, 1, None))  # plan-owned grid

    def capture(self):
        """Record one forward + adjoint into a HIP graph (torch CUDAGraph API)."""
        torch = self.torch
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.A._sync_stream()
            self.pair()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.A._sync_stream()
            self.pair()
        self.A._sync_stream()
        return self.graph.replay

    def stages(self):
        lib, p = self.lib, self.p
        return {
            "forward_grid": lambda: lib.mi355nufft_forward_grid(p, self.px, self.xs, self.pg, 1),
            "interp": lambda: lib.mi355nufft_interp(p, self.pg, self.py, 1),
            "spread": lambda: lib.mi355nufft_spread(p, self.py, self.pg, 1),
            "adjoint_image": lambda: lib.mi355nufft_adjoint_image(p, self.pg, self.pxo, self.xos, 1),
        }

    def time(self, fn, steps, warmup):
        torch = self.torch
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        s = torch.cuda.Event(enable_timing=True)
        e = torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(steps):
            fn()
        e.record()
        torch.cuda.synchronize()
        return s.elapsed_time(e) / steps

    def stage_times(self, steps, warmup):
        out = {}
        for name, fn in self.stages().items():
            def call(fn=fn):
                rc = fn()
                if rc:
                    self.L.check(rc)
            out[name] = self.time(call, steps, warmup)
        return out


def measure_atomic_peaks():
    import ctypes
    import torch
    from mrrt.nufft_amd import _lib
    lib = _lib.get_lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for kind, name in ((0, "lds_f64"), (1, "lds_u32"), (2, "global_f32"), (3, "lds_u64")):
        v = ctypes.c_double(0.0)
        _lib.check(lib.mi355nufft_measure_atomics(kind, ctypes.byref(v), stream))
        out[name] = v.value
    return out


def kernel_means(op, fn, n, atomics=None):
    """(gather_ms, scatter_ms): mean durations of the tiled kernels over n
    calls of fn, from the plan's dispatch-timestamped HIP events.  A dict
    passed as ``atomics`` receives the scatter's flush atomics per launch
    (global float atomic lane operations, counted in the kernel)."""
    op.set_kernel_timing(True)
    for _ in range(n):
        fn()
    _, g = op.read_kernel_timing("interp")
    _, sc = op.read_kernel_timing("spread")
    if atomics is not None:  # a separate run: the count slightly lengthens short launches
        atomics["per_launch"] = flush_atomics_per_launch(op, fn, n)
    op.set_kernel_timing(False)
    return g, sc


def flush_atomics_per_launch(op, fn, n):
    """Global float atomic lane operations of one scatter launch's tile
    flushes, counted in the kernel over n calls of fn."""
    op.set_kernel_timing(True, count_atomics=True)
    op.read_flush_atomics()
    for _ in range(n):
        fn()
    ns, _ = op.read_kernel_timing("spread")
    v = op.read_flush_atomics() / max(ns, 1)
    op.set_kernel_timing(False)
    return v


def flush_atomic_rate(atomics_per_launch, scatter_ms, peak):
    """The flush's global atomic lane operations per second over the whole
    scatter launch (a lower bound of the flush's own rate) and its fraction
    of the measured global float atomic rate."""
    rate = atomics_per_launch / (scatter_ms * 1e-3)
    return {"lane_ops_per_launch": int(atomics_per_launch), "lane_ops_per_s": round(rate, 1),
            "peak_lane_ops_per_s": round(peak, 1), "frac": round(rate / peak, 4),
            "peak_basis": "global_atomic_add_f32, coalesced, measured here (mi355nufft_measure_atomics kind 2)"}


def run_c4(steps=3, autotune=True, binning=None, loop=True):
    """C4: 24 frames x 32 coils, golden-angle frames of 13..89 spokes x 512,
    Nd 256^2, Kd 512^2, J 6, complex64, with the configuration's complex coil
    maps: SENSE encoding E x and E^H y of all frames (one image per frame)
    through one NufftFrames plan, captured in one HIP graph, against a
    per-frame loop of SenseOperator over plain NufftBase operators.
    ``binning`` = (bin_size, max_points, threads) fixes the binning instead
    of tuning it; ``loop=False`` skips the per-frame comparison."""
    import torch
    from mrrt.nufft_amd import NufftBase
    from mrrt.nufft_amd import workloads as W
    from mrrt.nufft_amd.batched import NufftFrames
    from mrrt.nufft_amd.mri import SenseOperator
    dev = torch.device("cuda", 0)
    wl = W.c4()
    F, C, Nd, Kd = len(wl["frames"]), wl["ncoils"], wl["Nd"], wl["Kd"]
    frames = wl["frames"]
    B = NufftFrames(Nd, frames, C, Jd=6, Kd=Kd, Ld=1024, device=dev)
    tuned = None
    if binning is not None:
        B.op.set_binning(binning[0], binning[1], "auto", binning[2])
        tuned = {"bin_size": binning[0], "max_points": binning[1], "threads": binning[2]}
    elif autotune:
        from mrrt.nufft_amd import autotune as at
        b, pmax, _ = at.tune(B.op, nreps=C, use_cache=False)
        info = B.op.plan_info()
        tuned = {"bin_size": b, "max_points": pmax, "threads": info["threads"]}
    smaps = torch.from_numpy(wl["smaps"].astype(np.complex64)).to(dev)
    E = SenseOperator(B, smaps)
    x = torch.from_numpy(wl["x"].astype(np.complex64)).to(dev)
    s = torch.cuda.Stream(device=dev)
    s.wait_stream(torch.cuda.current_stream(dev))
    with torch.cuda.stream(s):
        for _ in range(2):  # FFT set-up outside capture; leaves the plan grid zero
            E.adjoint_dev(E.forward_dev(x))
    torch.cuda.current_stream(dev).wait_stream(s)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        y_g = E.forward_dev(x)
        xa_g = E.adjoint_dev(y_g)
    B.op._sync_stream()
    t_b = _time_fn(graph.replay, steps, 1)
    t_eager = _time_fn(lambda: E.adjoint_dev(E.forward_dev(x)), steps, 1)
    t_l = float("nan")
    if loop:
        Es = [SenseOperator(NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev), smaps) for om in frames]

        def loop_frames():
            for f, Ef in enumerate(Es):
                Ef.adjoint_dev(Ef.forward_dev(x[f]))
        t_l = _time_fn(loop_frames, steps, 1)
        del Es
    # kernel durations of the frames transforms (768 coil images)
    xc = torch.randn((F, C) + Nd, dtype=torch.complex64, device=dev)
    yc = torch.empty((C, B.M), dtype=torch.complex64, device=dev)
    xac = torch.empty_like(xc)
    at4 = {}
    g_ms, s_ms = kernel_means(B.op, lambda: (B.forward(xc, out=yc), B.adjoint(yc, out=xac)), steps, at4)
    b_bytes = sum(interp_bytes(om.shape[0], 2, 512 * 512, C, 4) for om in frames)
    del graph
    return {"c4_frames_x_coils": [F, C], "c4_M_total": int(B.M),
            "c4_sense_graph_ms_per_pair": round(t_b, 3), "c4_sense_eager_ms_per_pair": round(t_eager, 3),
            "c4_sense_per_frame_loop_ms_per_pair": round(t_l, 3),
            "c4_speedup_graph_vs_loop": round(t_l / t_b, 3),
            "c4_transforms_per_s_graph": round(2 * F * C * 1000.0 / t_b, 1),
            "c4_interp_bytes_per_transform": int(b_bytes), "c4_autotuned": tuned,
            "c4_gather_kernel_ms": round(g_ms, 4), "c4_scatter_kernel_ms": round(s_ms, 4),
            "c4_scatter_flush_atomics_per_launch": int(at4["per_launch"]),
            "c4_gather_GBps": round(b_bytes / (g_ms * 1e-3) / 1e9, 1),
            "c4_scatter_GBps": round(b_bytes / (s_ms * 1e-3) / 1e9, 1),
            "c4_traffic_over_algorithmic": {k: round(v / b_bytes, 2) for k, v in TRAFFIC_BYTES_C4.items()},
            "c4_hbm_GBps_measured_traffic": {"interp": round(TRAFFIC_BYTES_C4["interp"] / (g_ms * 1e-3) / 1e9, 1),
                                             "spread": round(TRAFFIC_BYTES_C4["spread"] / (s_ms * 1e-3) / 1e9, 1)}}


def run_c5(steps=5):
    """C5: variable-density spiral 16 x 4096, 8 coils, Nd 256^2, c64; loss
    ||A(x, om) - y||^2 with y = A(x, om_pert) from the perturbed trajectory;
    forward + backward (grad x and grad omega) vs plain forward + adjoint;
    and SENSE with 8 coil maps and density weights."""
    import torch
    from mrrt.nufft_amd import NufftBase
    from mrrt.nufft_amd import trajectories as trj
    from mrrt.nufft_amd import workloads as W
    from mrrt.nufft_amd.autograd import nufft
    dev = torch.device("cuda", 0)
    wl = W.c5()
    Nd, Kd, C, om = wl["Nd"], wl["Kd"], wl["ncoils"], wl["om"]
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev)
    x = torch.from_numpy(wl["x"].astype(np.complex64)).to(dev).requires_grad_(True)
    omega = torch.from_numpy(om.astype(np.float32)).to(dev).requires_grad_(True)
    with torch.no_grad():
        y_t = nufft(op, x.detach(), torch.from_numpy(wl["om_pert"].astype(np.float32)).to(dev))
        _ = nufft(op, x.detach(), omega.detach())   # points back to om

    def ad_step():
        loss = ((nufft(op, x, omega) - y_t).abs() ** 2).sum()
        gx, gw = torch.autograd.grad(loss, (x, omega))
        return gx, gw

    xv = x.detach()

    def plain():
        y = op._forward_dev(xv.permute(1, 2, 0), C)
        op._adjoint_dev(y, C)

    t_ad = _time_fn(ad_step, steps, 2)
    t_pl = _time_fn(plain, steps, 2)
    out = {"c5_autograd_fwd_bwd_ms": round(t_ad, 4), "c5_plain_fwd_adj_ms": round(t_pl, 4),
           "c5_autograd_overhead": round(t_ad / t_pl, 3)}
    # SENSE encoding of one image with 8 coil maps and density weights:
    # E x, E^H W y, and the normal operator E^H W E x directly and through
    # the Toeplitz operator
    from mrrt.nufft_amd.mri import SenseOperator
    smaps = trj.coil_maps(C, Nd, seed=3).astype(np.complex64)
    w = np.random.default_rng(3).uniform(0.5, 1.0, om.shape[0]).astype(np.float32)
    E = SenseOperator(op, smaps, weights=w)
    xi = x.detach()[0].contiguous()
    t_fa = _time_fn(lambda: E.adjoint_dev(E.forward_dev(xi)), steps, 2)
    E.normal_dev(xi, toeplitz=True)
    t_tn = _time_fn(lambda: E.normal_dev(xi, toeplitz=True), steps, 2)
    out.update({"c5_sense_8coil_fwd_adj_ms": round(t_fa, 4), "c5_sense_8coil_normal_toeplitz_ms": round(t_tn, 4)})
    return out


def run_sparse(wl, tag, steps, warmup, fft):
    """Forward + adjoint with the precomputed sparse matrix (mode='sparse'):
    pair rate, stage times, set-up time and the matrix's device footprint
    (rows: values and each row's first taps; columns: values and sample
    indices)."""
    r = GpuRun(wl, mode="sparse")
    r.A.set_fft(fft)
    t = r.time(r.pair, steps, warmup)
    st = r.stage_times(max(5, steps // 2), min(warmup, 3))
    info = r.A.plan_info()
    vsize = 4 if wl["precision"] == "single" else 8
    out = {
        tag + "_sparse_pairs_per_s": round(1000.0 / t, 2),
        tag + "_sparse_stages_ms": {k: round(v, 4) for k, v in st.items()},
        tag + "_sparse_setup_s": round(r.setup_s, 3),
        tag + "_sparse_nnz": info["sparse_nnz"],
        tag + "_sparse_matrix_MB": round((info["sparse_nnz"] * (2 * vsize + 4) + r.M * r.ndim * 4) / 1e6, 1),
    }
    del r
    return out


def _time_fn(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1000.0 / steps
    return max(s.elapsed_time(e) / steps, wall)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class CpuPipeline:
    """The reference's table1 NUFFT on the host, stage by stage: its own C
    interpolation kernels (oracle/_ref, built from nufft_table.c with its
    -ffast-math -fopenmp flags; OpenMP over samples in the forward and over
    repetitions in the adjoint), the Python glue of nufft_forward / nufft_adj
    restated (mrrt/nufft/_nufft.py:1275-1397, 1459-1578), and scipy.fft for
    the grid FFT with as many workers as OpenMP threads (the reference's
    mrrt.utils.fftn is not in the tree).  Set-up (tables, sn, phases, tm) is
    excluded, as in the reference's own timings."""

    STAGES_F = ("deapod", "fft", "phase_before", "interp", "phase_after")
    STAGES_A = ("phase_after_conj", "scatter", "phase_before_conj", "ifft", "crop_scale_sn")

    def __init__(self, Nd, om, Kd, Jd, Ld, precision="single", workers=1):
        from oracle import nufft_oracle as O
        from oracle import ref_ctypes as R
        self.R = R
        self.kind = "reference" if R.available() else "port"
        self.B = O.NufftOracle(Nd, om, Jd=Jd, Kd=Kd, Ld=Ld, precision=precision)
        self.O = O
        self.workers = workers
        self.Nd, self.Kd = tuple(Nd), tuple(Kd)
        self.axes = tuple(range(len(Nd)))
        self.nk = int(np.prod(Kd))

    def _interp(self, xk):
        B = self.B
        if self.kind == "reference":
            return self.R.ref_interp_forward(xk, B.Kd, B.h, B.Jd, B.Ld, B.tm)
        return self.O.interp_forward(xk, B.Kd, B.h, B.Jd, B.Ld, B.tm)

    def _spread(self, y):
        B = self.B
        if self.kind == "reference":
            return self.R.ref_interp_adjoint(y, B.Kd, B.h, B.Jd, B.Ld, B.tm)
        return self.O.interp_adjoint(y, B.Kd, B.h, B.Jd, B.Ld, B.tm)

    def forward(self, x, t):
        import scipy.fft
        B = self.B
        nreps = x.shape[-1]
        c = time.perf_counter()
        xk = x * B.sn[..., None]
        t["deapod"] += time.perf_counter() - c
        c = time.perf_counter()
        xk = scipy.fft.fftn(xk, s=self.Kd, axes=self.axes, workers=self.workers)
        t["fft"] += time.perf_counter() - c
        c = time.perf_counter()
        xk = xk.astype(B.cdt, copy=False)
        xk *= B.phase_before[..., None]
        xk = xk.reshape((self.nk, nreps), order="F")
        t["phase_before"] += time.perf_counter() - c
        c = time.perf_counter()
        y = self._interp(xk)
        t["interp"] += time.perf_counter() - c
        c = time.perf_counter()
        y *= B.phase_after[:, None]
        t["phase_after"] += time.perf_counter() - c
        return y

    def adjoint(self, y, t):
        import scipy.fft
        B = self.B
        nreps = y.shape[-1]
        c = time.perf_counter()
        y = y * B.phase_after.conj()[:, None]
        t["phase_after_conj"] += time.perf_counter() - c
        c = time.perf_counter()
        g = self._spread(y)
        t["scatter"] += time.perf_counter() - c
        c = time.perf_counter()
        g = g.reshape(self.Kd + (nreps,), order="F")
        g *= B.phase_before.conj()[..., None]
        t["phase_before_conj"] += time.perf_counter() - c
        c = time.perf_counter()
        x = scipy.fft.ifftn(g, axes=self.axes, workers=self.workers)
        t["ifft"] += time.perf_counter() - c
        c = time.perf_counter()
        x = x[tuple(slice(n) for n in self.Nd)].astype(B.cdt, copy=False)
        x *= self.nk * B.adjoint_scalefactor
        x *= np.conj(B.sn)[..., None]
        t["crop_scale_sn"] += time.perf_counter() - c
        return x

    def pair(self, x):
        """One forward + adjoint; returns the per-stage seconds."""
        t = dict.fromkeys(self.STAGES_F + self.STAGES_A, 0.0)
        self.adjoint(self.forward(x, t), t)
        return t


def cpu_configs():
    """The CPU baseline's workloads: the benchmark configurations, each with
    its image batch (Fortran Nd + (nreps,))."""
    from mrrt.nufft_amd import trajectories as trj
    from mrrt.nufft_amd import workloads as W
    out = {}
    w = W.c1()
    out["c1"] = [(w, w["om"], w["x"].astype(np.complex64)[..., None])]
    w = W.c2("single")
    out["c2"] = [(w, w["om"], w["x"].astype(np.complex64)[..., None])]
    w = W.c2("double")
    out["c2_c128"] = [(w, w["om"], w["x"].astype(np.complex128)[..., None])]
    w = W.c3()
    out["c3"] = [(w, w["om"], w["x"].astype(np.complex64)[..., None])]
    w = W.c4()
    C = w["ncoils"]
    coil = np.moveaxis(w["smaps"][None] * w["x"][:, None], 1, -1).astype(np.complex64)  # (F, *Nd, C)
    out["c4"] = [(w, om, np.asfortranarray(coil[f])) for f, om in enumerate(w["frames"])]
    w = W.c5()
    out["c5"] = [(w, w["om"], np.asfortranarray(np.moveaxis(w["x"], 0, -1).astype(np.complex64)))]
    return out


def cpu_run_config(name, items, runs, warmup=1):
    """Median over ``runs`` of one forward + adjoint of every item (the
    per-frame loop for C4); per-stage medians in ms."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    pipes = [(CpuPipeline(w["Nd"], om, w["Kd"], w["Jd"], w["Ld"], w["precision"], workers=threads), x)
             for (w, om, x) in items]
    for _ in range(warmup):
        for P, x in pipes:
            P.pair(x)
    tot, per = [], []
    for _ in range(runs):
        t0 = time.perf_counter()
        acc = dict.fromkeys(CpuPipeline.STAGES_F + CpuPipeline.STAGES_A, 0.0)
        for P, x in pipes:
            for k, v in P.pair(x).items():
                acc[k] += v
        tot.append(time.perf_counter() - t0)
        per.append(acc)
    med = float(np.median(tot))
    stages = {k: round(1e3 * float(np.median([p[k] for p in per])), 3) for k in per[0]}
    return {"config": name, "threads": threads, "fft_workers": threads, "kind": pipes[0][0].kind,
            "runs": runs, "ms_per_pair": round(1e3 * med, 3), "pairs_per_s": round(1.0 / med, 4),
            "stages_ms": stages, "cpu": cpu_model()}


def cpu_baseline(seconds):
    """The reference pipeline (CpuPipeline) on the GPU box's host cores, C2
    complex64, for about ``seconds``."""
    from mrrt.nufft_amd import workloads as W
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    w = W.c2("single")
    P = CpuPipeline(w["Nd"], w["om"], w["Kd"], w["Jd"], w["Ld"], "single", workers=threads)
    x = w["x"].astype(np.complex64)[..., None]
    P.pair(x)
    t0 = time.perf_counter()
    n = 0
    while True:
        P.pair(x)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = (time.perf_counter() - t0) / n
    return {"value": round(1.0 / dt, 4), "unit": "pairs/s", "cores": threads, "kind": P.kind,
            "sample": "%d forward+adjoint pairs of the 2-D 256^2 / 512^2 golden-angle workload (complex64) on the "
                      "GPU box's host (%s, %d threads): the reference's C table1 kernels (OpenMP over samples "
                      "forward, one thread per coil adjoint) + scipy.fft with %d workers" % (n, cpu_model(), threads,
                                                                                              threads)}


def main_cpu(args):
    cfgs = cpu_configs()
    for name in args.cpu_configs.split(","):
        print("bench: cpu %s (%d runs, %s threads)" % (name, args.cpu_runs, os.environ.get("OMP_NUM_THREADS", "1")),
              file=sys.stderr, flush=True)
        print(json.dumps(cpu_run_config(name, cfgs[name], args.cpu_runs)), flush=True)


def main():
    args = parse()
    if args.cpu_only:
        main_cpu(args)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: every run here uses a single MI355X "
              "(the path shards by frames and coils, or by slab; only one GPU is measured)", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)

    if args.calibrate_copy:
        a = torch.empty(64 << 20, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        print("bench: calibration copies done (3 x 256 MiB)", file=sys.stderr, flush=True)

    wl = c2_workload("single")
    run = GpuRun(wl)
    # grid FFT of the plan: the fused pruned passes or rocFFT, whichever
    # times faster here for this geometry (autotune.tune_fft)
    from mrrt.nufft_amd import autotune
    if args.fft == "tuned":
        fft_choice, fft_ms = autotune.tune_fft(run.A, nreps=1)
    else:
        fft_choice, fft_ms = ("auto" if args.fft == "fused" else "rocfft"), {}
        run.A.set_fft(fft_choice)
    print("bench: plan ready (M=%d, fft %s %s); timing %d steps after %d warmup"
          % (run.M, fft_choice, json.dumps({k: round(v, 5) for k, v in fft_ms.items()}), args.steps, args.warmup),
          file=sys.stderr, flush=True)
    ms = _time_fn(run.pair, args.steps, args.warmup)
    ms_graph = _time_fn(run.capture(), args.steps, args.warmup)
    if args.graph:
        ms, ms_graph = ms_graph, ms
    stages = run.stage_times(max(10, args.steps // 2), 3)
    print("bench: stages (ms) %s" % json.dumps({k: round(v, 5) for k, v in stages.items()}),
          file=sys.stderr, flush=True)
    # Kernel durations: a second eager run of the same steps in which every
    # gather / scatter launch carries a HIP event pair timestamped by its
    # dispatch on the plan's stream (the kernel's execution only).
    run.A.set_kernel_timing(True)
    _time_fn(run.pair, args.steps, 0)
    run.A.set_kernel_timing(False)
    kern = {w: run.A.read_kernel_timing(w) for w in ("interp", "spread")}
    flush_atomics = flush_atomics_per_launch(run.A, run.pair, 10)
    print("bench: kernel means (ms) %s" % json.dumps({k: [v[0], round(v[1], 5)] for k, v in kern.items()}),
          file=sys.stderr, flush=True)
    b = interp_bytes(run.M, 2, run.nk, 1, 4)
    dom = "interp" if kern["interp"][1] >= kern["spread"][1] else "spread"
    achieved = b / (kern[dom][1] * 1e-3) / 1e9
    traffic = TRAFFIC_BYTES.get(dom)
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                "kernel": KERNEL_NAMES[dom], "kernel_ms": round(kern[dom][1], 5), "launches": kern[dom][0],
                "bytes_per_launch": b,
                "other_kernel": {"kernel": KERNEL_NAMES["interp" if dom == "spread" else "spread"],
                                 "kernel_ms": round(kern["interp" if dom == "spread" else "spread"][1], 5)}}
    # the scatter's atomic roofs, measured on this GPU: LDS float64 and
    # 32-bit integer atomic adds at scattered addresses, and coalesced global
    # float32 atomic adds (mi355nufft_measure_atomics)
    peaks = measure_atomic_peaks()
    print("bench: atomic peaks (lane ops/s) %s" % json.dumps({k: "%.3g" % v for k, v in peaks.items()}),
          file=sys.stderr, flush=True)
    # M * J^d complex accumulations per scatter launch, two LDS atomic lane
    # operations each: ds_add_u32 with the fixed-point tile, else ds_add_f64
    acc = run.M * 6 ** 2
    t_sc = kern["spread"][1] * 1e-3
    fxp = run.A.plan_info()["fixed_point_scatter"]
    lds_peak = peaks["lds_u32"] if fxp else peaks["lds_f64"]
    roofline["scatter_lds_atomics"] = {
        "accumulations_per_s": round(acc / t_sc, 1),
        "lane_atomics_per_s": round(2 * acc / t_sc, 1),
        "peak_lane_atomics_per_s": round(lds_peak, 1),
        "frac": round(2 * acc / t_sc / lds_peak, 4),
        "peak_basis": ("ds_add_u32" if fxp else "ds_add_f64") + " at scattered addresses, measured here "
                      "(mi355nufft_measure_atomics kind %d)" % (1 if fxp else 0),
        "accumulation": "fixed32" if fxp else "float64",
    }
    roofline["atomic_peaks_lane_ops_per_s"] = {k: round(v, 1) for k, v in peaks.items()}
    # the flush's global float atomics (counted in the kernel) against the
    # measured global float atomic rate
    roofline["scatter_flush_global_atomics"] = flush_atomic_rate(flush_atomics, kern["spread"][1], peaks["global_f32"])
    extra = {}
    if not args.no_extra:
        other = "rocfft" if fft_choice == "auto" else "auto"
        run.A.set_fft(other)
        t_other = _time_fn(run.pair, args.steps, args.warmup)
        run.A.set_fft(fft_choice)
        extra["c2_pairs_per_s_with_fft_" + ("fused" if other == "auto" else "rocfft")] = round(1000.0 / t_other, 2)
        # the Python drop-in on host (numpy) arrays: PCIe copies included
        xh = np.asfortranarray(wl["x"].astype(np.complex64))
        for _ in range(5):
            yh = run.A.fft(xh)
            run.A.adj(yh)
        nh = max(50, args.steps)
        t0 = time.perf_counter()
        for _ in range(nh):
            yh = run.A.fft(xh)
            run.A.adj(yh)
        extra["c2_host_arrays_pairs_per_s"] = round(nh / (time.perf_counter() - t0), 2)
        extra["c2_host_arrays_pairs_timed"] = nh
        # Toeplitz normal operator A^H A on the same points (circulant 512^2)
        from mrrt.nufft_amd.toeplitz import ToeplitzNormal
        tn = ToeplitzNormal(run.A)
        xt = run.x
        t_n = _time_fn(lambda: tn.apply_dev(xt), args.steps, args.warmup)
        extra["c2_toeplitz_normal_per_s"] = round(1000.0 / t_n, 2)
        extra["c2_toeplitz_vs_pair_speedup"] = round(ms / t_n, 3)
        del tn
        r2 = GpuRun(c2_workload("double"))
        t2 = r2.time(r2.pair, max(5, args.steps // 2), 3)
        extra["c2_complex128_pairs_per_s"] = round(1000.0 / t2, 2)
        del r2
        # sparse-matrix mode (mode='sparse') on the same workload
        print("bench: sparse mode", file=sys.stderr, flush=True)
        extra.update(run_sparse(c2_workload("single"), "c2", args.steps, args.warmup, fft_choice))
        torch.cuda.empty_cache()
        r3 = GpuRun(c3_workload())
        t3 = r3.time(r3.pair, 5, 1)
        st3 = r3.stage_times(5, 1)
        at3 = {}
        g3, s3 = kernel_means(r3.A, r3.pair, 5, at3)
        extra["c3_3d_complex64_pairs_per_s"] = round(1000.0 / t3, 3)
        extra["c3_stages_ms"] = {k: round(v, 4) for k, v in st3.items()}
        b3 = interp_bytes(r3.M, 3, r3.nk, 1, 4)
        extra["c3_gather_kernel_ms"] = round(g3, 4)
        extra["c3_scatter_kernel_ms"] = round(s3, 4)
        extra["c3_gather_GBps"] = round(b3 / (g3 * 1e-3) / 1e9, 1)
        extra["c3_scatter_GBps"] = round(b3 / (s3 * 1e-3) / 1e9, 1)
        extra["c3_traffic_over_algorithmic"] = {k: round(v / b3, 2) for k, v in TRAFFIC_BYTES_C3.items()}
        extra["c3_hbm_GBps_measured_traffic"] = {"interp": round(TRAFFIC_BYTES_C3["interp"] / (g3 * 1e-3) / 1e9, 1),
                                                 "spread": round(TRAFFIC_BYTES_C3["spread"] / (s3 * 1e-3) / 1e9, 1)}
        info3 = r3.A.plan_info()
        acc3 = r3.M * 6 ** 3
        # the 3-D fixed-point tile makes one ds_add_u64 per tap (packed
        # components), the float64 tile two ds_add_f64
        if info3["fixed_point_scatter"]:
            extra["c3_scatter_lds_atomic_frac"] = round(acc3 / (s3 * 1e-3) / peaks["lds_u64"], 4)
        else:
            extra["c3_scatter_lds_atomic_frac"] = round(2 * acc3 / (s3 * 1e-3) / peaks["lds_f64"], 4)
        extra["c3_scatter_flush_global_atomics"] = flush_atomic_rate(at3["per_launch"], s3, peaks["global_f32"])
        extra["c3_scatter_accumulation"] = "fixed32" if info3["fixed_point_scatter"] else "float64"
        del r3
        torch.cuda.empty_cache()
        extra.update(run_sparse(c3_workload(), "c3", 5, 1, "auto"))
        torch.cuda.empty_cache()
        print("bench: frames batch (24 x 32)", file=sys.stderr, flush=True)
        c4 = run_c4()
        c4["c4_scatter_flush_global_atomics"] = flush_atomic_rate(
            c4["c4_scatter_flush_atomics_per_launch"], c4["c4_scatter_kernel_ms"], peaks["global_f32"])
        extra.update(c4)
        torch.cuda.empty_cache()
        print("bench: autograd spiral", file=sys.stderr, flush=True)
        extra.update(run_c5(steps=20))
    cpu = None
    if not args.no_cpu_baseline:
        print("bench: cpu baseline (~%.0f s)" % args.cpu_seconds, file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_seconds)
    value = 1000.0 / ms
    line = {
        "metric": "NUFFT fwd+adj pairs/s (2D 256^2, Kd 512^2, J6, L1024, 205,824 golden-angle samples)",
        "value": round(value, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": "2D NUFFT table1 Nd=(256,256) Kd=(512,512) Jd=6 Ld=1024 golden-angle 402x512 "
                               "complex64, forward+adjoint incl. the grid FFTs", "M": run.M, "nreps": 1,
                   "fft": "fused pruned passes" if fft_choice == "auto" else "rocFFT"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": {k: round(v, 5) for k, v in stages.items()},
        "step": "HIP graph replay of forward+adjoint" if args.graph else "eager launches",
        "ms_per_step_other": round(ms_graph, 5),
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
