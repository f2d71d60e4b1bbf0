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

# float64 LDS atomic adds per second over the chip: 64 lanes per 9.06 CU
# cycles (measured for ds_add_f64 without bank conflicts), 256 CUs, 2.4 GHz
LDS_F64_ATOMIC_PEAK_LANES_PER_S = 64 / 9.06 * 256 * 2.4e9

KERNEL_NAMES = {"interp": "interp_tiled_kernel<float, 2, false, 6, true>",
                "spread": "spread_tiled_kernel<float, 2, false, 6, true>"}

# HBM bytes per launch of the 2-D 256^2 complex64 kernels, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes over this bench (--calibrate-copy), each
# scaled by what the counter reads for a 256 MiB device copy (FETCH_SIZE
# x 2.0, WRITE_SIZE x 1.0); profiles/nufft2d_256_c64_hbm_traffic.md.
TRAFFIC_BYTES = {"spread": 20200960, "interp": 13982822}


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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run three 256 MiB device-to-device copies (a known-size reference for "
                         "FETCH_SIZE / WRITE_SIZE counter passes)")
    return ap.parse_args()


def c2_workload(precision="single"):
    from mrrt.nufft_amd import trajectories as trj
    Nd, Kd = (256, 256), (512, 512)
    om = trj.golden_angle_2d(402, 512)
    x = trj.complex_gaussian(Nd, seed=0) + trj.ellipse_phantom(Nd, 10, seed=0)
    return dict(Nd=Nd, Kd=Kd, Jd=6, Ld=1024, om=om, x=x, precision=precision)


def c3_workload():
    from mrrt.nufft_amd import trajectories as trj
    Nd, Kd = (128, 128, 128), (256, 256, 256)
    om = trj.kooshball_3d(8192, 256, seed=0)
    x = trj.complex_gaussian(Nd, seed=0)
    return dict(Nd=Nd, Kd=Kd, Jd=6, Ld=512, om=om, x=x, precision="single")


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


def run_c4(steps=3, autotune=True, binning=None, loop=True):
    """24 frames x 32 coils, golden-angle frames with 13..89 spokes x 512,
    Nd 256^2, Kd 512^2, J 6, complex64: one graph-captured frames plan vs a
    per-frame loop of plain operators (32 coils each).  ``binning`` =
    (bin_size, max_points, threads) fixes the binning instead of tuning it;
    ``loop=False`` skips the per-frame comparison (e.g. for profiling)."""
    import torch
    from mrrt.nufft_amd import NufftBase
    from mrrt.nufft_amd import trajectories as trj
    from mrrt.nufft_amd.batched import NufftFrames
    dev = torch.device("cuda", 0)
    F, C, Nd, Kd = 24, 32, (256, 256), (512, 512)
    frames = trj.golden_angle_frames(F, 512, 13, 89, seed=0)
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
    rng = torch.Generator(device=dev).manual_seed(0)
    x = torch.randn((F, C) + Nd, dtype=torch.complex64, device=dev, generator=rng)
    y = torch.empty((C, B.M), dtype=torch.complex64, device=dev)
    xa = torch.empty_like(x)
    B.capture(x, y, xa)

    def batched():
        B.replay_forward()
        B.replay_adjoint()

    ops = [NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev, order="C") for om in frames] if loop else []

    def loop_frames():
        for f, op in enumerate(ops):
            xv = x[f].permute(1, 2, 0)
            yf = op._forward_dev(xv, C)
            op._adjoint_dev(yf, C)

    do_loop = loop
    t_b = _time_fn(batched, steps, 1)
    t_l = _time_fn(loop_frames, steps, 1) if do_loop else float("nan")
    # stage split of one batched forward + adjoint (plan stages, 768 images)
    import ctypes
    from mrrt.nufft_amd import _lib
    op, lib, vp = B.op, B.op._lib, ctypes.c_void_p
    xv = B._image_view(x)
    xs = _lib.i64arr(list(xv.stride()))
    g = B._grid
    nimg = F * C
    op._sync_stream()
    stages = {
        "forward_grid": lambda: _lib.check(lib.mi355nufft_forward_grid(op._plan, vp(xv.data_ptr()), xs,
                                                                       vp(g.data_ptr()), nimg)),
        "interp": lambda: _lib.check(lib.mi355nufft_interp(op._plan, vp(g.data_ptr()), vp(y.data_ptr()), C)),
        "spread": lambda: _lib.check(lib.mi355nufft_spread(op._plan, vp(y.data_ptr()), vp(g.data_ptr()), C)),
        "adjoint_image": lambda: _lib.check(lib.mi355nufft_adjoint_image(op._plan, vp(g.data_ptr()),
                                                                         vp(xv.data_ptr()), xs, nimg)),
    }
    st = {k: round(_time_fn(fn, steps, 1), 4) for k, fn in stages.items()}
    b_bytes = sum(interp_bytes(om.shape[0], 2, 512 * 512, C, 4) for om in frames)
    return {"c4_frames_x_coils": [F, C], "c4_M_total": int(B.M),
            "c4_batched_graph_ms_per_pair": round(t_b, 3), "c4_per_frame_loop_ms_per_pair": round(t_l, 3),
            "c4_speedup_batched_vs_loop": round(t_l / t_b, 3),
            "c4_transforms_per_s_batched": round(2 * F * C * 1000.0 / t_b, 1),
            "c4_interp_bytes_per_transform": int(b_bytes), "c4_autotuned": tuned, "c4_stages_ms": st,
            "c4_interp_GBps": round(b_bytes / (st["interp"] * 1e-3) / 1e9, 1),
            "c4_spread_GBps": round(b_bytes / (st["spread"] * 1e-3) / 1e9, 1)}


def run_c5(steps=5):
    """Differentiable spiral: 16 x 4096 samples, 8 coils, Nd 256^2, c64;
    forward + backward (grad x and grad omega) vs plain forward + adjoint."""
    import torch
    from mrrt.nufft_amd import NufftBase
    from mrrt.nufft_amd import trajectories as trj
    from mrrt.nufft_amd.autograd import nufft
    dev = torch.device("cuda", 0)
    Nd, Kd, C = (256, 256), (512, 512), 8
    om = trj.spiral_2d(16, 4096, seed=0)
    op = NufftBase(Nd, om, Jd=6, Kd=Kd, Ld=1024, device=dev)
    rng = torch.Generator(device=dev).manual_seed(1)
    x = torch.randn((C,) + Nd, dtype=torch.complex64, device=dev, generator=rng).requires_grad_(True)
    omega = torch.from_numpy(om.astype(np.float32)).to(dev).requires_grad_(True)
    y_t = torch.randn((C, om.shape[0]), dtype=torch.complex64, device=dev, generator=rng)

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
    g = np.meshgrid(np.linspace(-1, 1, Nd[0]), np.linspace(-1, 1, Nd[1]), indexing="ij")
    cen = np.random.default_rng(2).uniform(-1, 1, (C, 2))
    smaps = np.stack([np.exp(-((g[0] - a) ** 2 + (g[1] - b) ** 2)) for a, b in cen]).astype(np.complex64)
    w = np.random.default_rng(3).uniform(0.5, 1.0, om.shape[0]).astype(np.float32)
    E = SenseOperator(op, smaps, weights=w)
    xi = x.detach()[0].contiguous()
    ys = E.forward_dev(xi)
    t_fa = _time_fn(lambda: E.adjoint_dev(E.forward_dev(xi)), steps, 2)
    E.normal_dev(xi, toeplitz=True)
    t_tn = _time_fn(lambda: E.normal_dev(xi, toeplitz=True), steps, 2)
    del ys
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


def cpu_baseline(seconds):
    """Reference C kernels (oracle/_ref) + numpy FFT on the host, C2 c64."""
    from oracle import nufft_oracle as O
    from oracle import ref_ctypes as R
    wl = c2_workload("single")
    kind = "reference" if R.available() else "port"
    interp = (R.ref_interp_forward, R.ref_interp_adjoint) if R.available() else "numpy"
    B = O.NufftOracle(wl["Nd"], wl["om"], Jd=6, Kd=wl["Kd"], Ld=1024, precision="single", interp=interp)
    x = wl["x"].astype(np.complex64)
    y = B.fft(x)
    B.adj(y)
    t0 = time.perf_counter()
    n = 0
    while True:
        y = B.fft(x)
        B.adj(y)
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = (time.perf_counter() - t0) / n
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": round(1.0 / dt, 4), "unit": "pairs/s", "cores": threads, "kind": kind,
            "sample": "%d forward+adjoint pairs of the same 2-D 256^2 workload (complex64); reference C "
                      "table1 kernels (OpenMP forward, 1-thread-per-coil adjoint) + numpy FFT" % n}


def main():
    args = parse()
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
    # the scatter's LDS-atomic view: M * J^d complex accumulations per launch,
    # two float64 LDS atomic adds each, against the LDS atomic peak
    acc = run.M * 6 ** 2
    t_sc = kern["spread"][1] * 1e-3
    lane_peak = LDS_F64_ATOMIC_PEAK_LANES_PER_S
    roofline["scatter_lds_atomics"] = {
        "accumulations_per_s": round(acc / t_sc, 1),
        "lane_atomics_per_s": round(2 * acc / t_sc, 1),
        "peak_lane_atomics_per_s": round(lane_peak, 1),
        "frac": round(2 * acc / t_sc / lane_peak, 4),
        "peak_basis": "ds_add_f64: 64 lanes per 9.06 CU-cycles (measured, conflict-free), 256 CUs, 2.4 GHz",
    }
    extra = {}
    if not args.no_extra:
        other = "rocfft" if fft_choice == "auto" else "auto"
        run.A.set_fft(other)
        t_other = _time_fn(run.pair, args.steps, args.warmup)
        run.A.set_fft(fft_choice)
        extra["c2_pairs_per_s_with_fft_" + ("fused" if other == "auto" else "rocfft")] = round(1000.0 / t_other, 2)
        # the Python drop-in on host (numpy) arrays: PCIe copies included
        xh = np.asfortranarray(wl["x"].astype(np.complex64))
        yh = run.A.fft(xh)
        run.A.adj(yh)
        nh = max(5, args.steps // 5)
        t0 = time.perf_counter()
        for _ in range(nh):
            yh = run.A.fft(xh)
            run.A.adj(yh)
        extra["c2_host_arrays_pairs_per_s"] = round(nh / (time.perf_counter() - t0), 2)
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
        extra["c3_3d_complex64_pairs_per_s"] = round(1000.0 / t3, 3)
        extra["c3_stages_ms"] = {k: round(v, 4) for k, v in st3.items()}
        b3 = interp_bytes(r3.M, 3, r3.nk, 1, 4)
        extra["c3_interp_GBps"] = round(b3 / (st3["interp"] * 1e-3) / 1e9, 1)
        extra["c3_spread_GBps"] = round(b3 / (st3["spread"] * 1e-3) / 1e9, 1)
        del r3
        torch.cuda.empty_cache()
        extra.update(run_sparse(c3_workload(), "c3", 5, 1, "auto"))
        torch.cuda.empty_cache()
        print("bench: frames batch (24 x 32)", file=sys.stderr, flush=True)
        extra.update(run_c4())
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
