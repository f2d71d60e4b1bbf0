"""Autotuner for the tiled interpolation kernels.

Picks the bin size, the points per subproblem and the workgroup size of the
tiled kernels of a plan by
timing the gather and the scatter stages on the plan's own points with HIP
events, and caches the choice per geometry in a JSON file
(``$MI355NUFFT_TUNE_CACHE`` or ``~/.cache/mi355nufft/autotune.json``).
The cache key holds the library's ABI version, dimensions, image size,
grid, kernel width, table size, precision, phasing, number of points
(rounded to a power of two), repetitions, frames and the scatter's
accumulation; trajectories of the same kind and size share an entry.  An
entry the library rejects is dropped and tuned again.
"""

import ctypes
import json
import math
import os

import numpy as np
import torch

from . import _lib

__all__ = ["tune", "tune_fft", "candidates", "cache_path"]

# (bin size, points per subproblem, threads per workgroup)
_C2 = [((b, b), p, t) for b in (16, 32) for p in (64, 128, 256, 384, 512) for t in (64, 128, 256) if p >= t] + \
    [(b, p, 256) for b in ((64, 32), (32, 64), (64, 64)) for p in (512, 1024)]
_C3 = [(b, p, t) for b in ((8, 8, 8), (16, 8, 8), (16, 16, 16)) for p in (512, 1024, 2048) for t in (64, 256)] + \
    [((16, 16, 16), p, 256) for p in (2560, 3072, 4096)]
_C1 = [((b,), p, t) for b in (256, 512, 1024) for p in (256, 512) for t in (64, 256)]


def candidates(ndim):
    return {1: _C1, 2: _C2, 3: _C3}[ndim]


def cache_path():
    p = os.environ.get("MI355NUFFT_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "mi355nufft", "autotune.json")


def _key(op, nreps):
    mb = 1 << int(round(math.log2(max(op.M, 1))))
    abi = _lib.get_lib().mi355nufft_abi_version()
    fxp = int(op.plan_info()["fixed_point_scatter"])
    return "|".join(str(v) for v in (abi, op.ndim, op.Nd, op.Kd, op.Jd, op.Ld, op.precision, op.phasing, mb, nreps,
                                     getattr(op, "nframes", 1), fxp))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass


def _time(fn, reps):
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    fn()
    best = float("inf")
    for _ in range(reps):
        s.record()
        fn()
        e.record()
        e.synchronize()
        best = min(best, s.elapsed_time(e))
    return best


def tune_fft(op, nreps=1, pairs=20, rounds=3):
    """Choose the grid FFT of ``op`` ("auto": pruned passes fused with the
    deapodization, or "rocfft") by timing back-to-back forward + adjoint
    pairs with each (``rounds`` alternating rounds of ``pairs`` pairs, best
    round per path), and leave the plan set to the faster: the fused passes
    unless rocFFT is more than 2% faster.  Needs the points set; returns
    (choice, {choice: ms per pair})."""
    op.set_fft("auto")
    if not op.plan_info()["fused_fft"]:  # no fused passes for this grid: auto is rocFFT
        return "auto", {}
    x = torch.randn(tuple(op.Nd) + (nreps,), dtype=op._tcdt, device=op.device)

    def run():
        for _ in range(pairs):
            op._adjoint_dev(op._forward_dev(x, nreps), nreps)
    out = {}
    for _ in range(rounds):
        for m in ("auto", "rocfft"):
            op.set_fft(m)
            t = _time(run, 1) / pairs
            out[m] = min(out.get(m, float("inf")), t)
    # rocFFT only when it is faster beyond the timing noise (2%)
    best = "rocfft" if out["rocfft"] < 0.98 * out["auto"] else "auto"
    op.set_fft(best)
    return best, out


def tune(op, nreps=1, cands=None, reps=3, use_cache=True, verbose=False):
    """Choose bin size / points per subproblem for ``op`` (a NufftBase, or
    the ``.op`` of a NufftFrames) with ``nreps`` repetitions per frame.

    Returns ``(bin_size, max_points, timings)``; the plan is left configured
    with the choice.  ``timings`` maps each candidate to (gather_ms, scatter_ms).
    """
    key = _key(op, nreps)
    table = _load() if use_cache else {}
    if use_cache and key in table:
        b, p = table[key]["bin_size"], table[key]["max_points"]
        try:
            op.set_binning(tuple(b), p, "auto", table[key].get("threads", 0))
            op.set_fft(table[key].get("fft", "auto"))
            return tuple(b), p, table[key].get("timings", {})
        except (ValueError, KeyError, TypeError):
            del table[key]  # a stale entry this library rejects: tune again
    cands = cands or candidates(op.ndim)
    nimg = getattr(op, "nframes", 1) * nreps
    nk = int(np.prod(op.Kd))
    g = torch.randn((nimg, nk), dtype=op._tcdt, device=op.device)
    y = torch.randn((nreps, op.M), dtype=op._tcdt, device=op.device)
    lib = op._lib
    vp = ctypes.c_void_p
    timings = {}
    best = None
    mode = "tiled_atomic"
    for b, p, th in cands:
        try:
            op.set_binning(b, p, mode, th)
        except ValueError:
            continue
        op._sync_stream()
        tf = _time(lambda: _lib.check(lib.mi355nufft_interp(op._plan, vp(g.data_ptr()), vp(y.data_ptr()), nreps)),
                   reps)
        ts = _time(lambda: _lib.check(lib.mi355nufft_spread(op._plan, vp(y.data_ptr()), vp(g.data_ptr()), nreps)),
                   reps)
        timings[str((b, p, th))] = (tf, ts)
        if verbose:
            print("autotune", b, p, th, "gather %.4f ms scatter %.4f ms" % (tf, ts))
        if best is None or tf + ts < best[0]:
            best = (tf + ts, b, p, th)
    if best is None:
        op.set_binning(None, 0, "auto")
        return None, 0, timings
    _, b, p, th = best
    op.set_binning(b, p, "auto", th)
    fft, fft_ms = ("auto", {}) if getattr(op, "nframes", 1) > 1 else tune_fft(op, nreps)
    if verbose:
        print("autotune fft", fft, fft_ms)
    if use_cache:
        table[key] = {"bin_size": list(b), "max_points": p, "threads": th, "timings": timings, "fft": fft}
        _store(table)
    return tuple(b), p, timings
