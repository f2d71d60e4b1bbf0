"""NufftBase drop-in: the table1 NUFFT on MI355X.

Keeps the constructor and call surface of the reference's ``NufftBase``
(mrrt/nufft/_nufft.py:216-482): ``NufftBase(Nd, omega, Jd, Kd, precision,
mode, Ld, ortho, n_shift, phasing, adjoint_scalefactor, preplan_cufft, order,
verbose, on_gpu)`` with ``.fft(x)`` and ``.adj(k)``, Fortran-ordered
semantics (image ``Nd + (nreps,)``, samples ``(M, nreps)``) and the
``order='C'`` repetition swap (:415-423).

All transform work runs in ``libmi355nufft.so`` (HIP kernels + rocFFT) on the
current torch stream; torch tensors serve as device buffers.  Inputs may be
numpy arrays (copied to the device, results returned as numpy, like the
reference's CPU path) or torch tensors (results stay on the device).
"""

import ctypes
import threading
import warnings
from math import sqrt

import numpy as np
import torch

from . import _lib
from ._kernel_math import beatty_alpha, build_table, deapod_factors, kaiser_bessel

__all__ = ["NufftBase", "nufft_forward", "nufft_adj"]

_SUPPORTED_REAL = (np.float32, np.float64)


def _as_tuple(seq, type=int, n=None):
    """_utils.py:130-137."""
    if np.isscalar(seq):
        if n is None:
            raise ValueError("for scalar, seq, n must be specified")
        return (type(seq),) * n
    elif n is not None and len(seq) != n:
        raise ValueError("array did not have the expected size of {}".format(n))
    return tuple(type(s) for s in seq)


def _default_device(device):
    if device is not None:
        return torch.device(device)
    if not torch.cuda.is_available():
        raise RuntimeError("mrrt.nufft_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def _f_reshape(t, shape):
    """numpy ``reshape(shape, order='F')`` for a torch tensor."""
    shape = tuple(shape)
    if tuple(t.shape) == shape:
        return t
    if tuple(t.shape) + (1,) == shape:
        return t.unsqueeze(-1)
    flat = t.permute(*reversed(range(t.dim()))).reshape(-1)
    return flat.reshape(tuple(reversed(shape))).permute(*reversed(range(len(shape))))


class NufftBase(object):
    """Table-mode ("table1") or sparse-matrix NUFFT operator on one MI355X.

    Parameters follow the reference (mrrt/nufft/_nufft.py:122-233).  ``om``
    is accepted as an alias of ``omega``.  ``mode`` is ``'sparse'`` (the
    precomputed interpolation matrix of :751-877, held on the device) or
    contains "table" (``'table'`` or ``'table1'``).  ``device`` selects the
    GPU (default: the current torch device).
    """

    def __init__(self, Nd, omega=None, Jd=4, Kd=None, precision="single", mode="table", Ld=1024,
                 ortho=False, n_shift=None, phasing="real", adjoint_scalefactor=1.0,
                 preplan_cufft=True, order="F", verbose=False, on_gpu=True, om=None, device=None):
        if omega is None:
            omega = om
        if omega is None:
            raise ValueError("omega is required")
        self.verbose = verbose
        self.order = order
        if order not in ("F", "C"):
            raise ValueError("order must be 'F' or 'C'")
        self.on_gpu = True
        self.device = _default_device(device)
        if np.isscalar(Nd):
            Nd = (Nd,)
        self._Nd = _as_tuple(Nd, type=int)
        self.ndim = len(self._Nd)
        if phasing not in ("real", "complex"):
            raise ValueError("Invalid phasing: {}. phasing must be 'real' or 'complex'".format(phasing))
        self._phasing = phasing
        self._Jd = _as_tuple(Jd, type=int, n=self.ndim)
        if Kd is None:
            Kd = tuple(int(1.5 * n) for n in self._Nd)
        self._Kd = _as_tuple(Kd, type=int, n=self.ndim)
        if self.ndim > 3:
            raise NotImplementedError("dimensions > 3 not implemented")
        self.ortho = ortho
        self.scale_ortho = sqrt(float(np.prod(self._Kd))) if ortho else 1
        if mode == "sparse":                            # _nufft.py:321-350
            self._sparse = True
        elif "table" in mode:
            self._sparse = False
        elif mode == "exact":
            raise ValueError("mode exact not implemented")
        else:
            raise ValueError("Invalid NUFFT mode: {}".format(mode))
        self._mode = mode
        self._Ld = int(Ld) if Ld is not None else 2 ** 10
        self.adjoint_scalefactor = adjoint_scalefactor
        self.preplan_cufft = preplan_cufft

        omega_np, omega_t = self._check_omega(omega)
        if precision == "auto":
            precision = "single" if omega_t.dtype == torch.float32 else "double"
        self._set_dtypes(precision)

        if n_shift is None:
            self._n_shift = (0.0,) * self.ndim
        else:
            self._n_shift = _as_tuple(n_shift, type=float, n=self.ndim)
        if (self.ndim != len(self._Jd)) or (self.ndim != len(self._Kd)):
            raise ValueError("Inconsistent Dimensions")
        if phasing == "real":
            self.n_mid = tuple(n // 2 for n in self._Nd)
        else:
            self.n_mid = tuple((n - 1) / 2.0 for n in self._Nd)
        for d in range(self.ndim):
            if self._Nd[d] > self._Kd[d]:
                raise ValueError("Kd must be >= Nd on every axis")
        odd_L = self._Ld % 2 == 1
        if odd_L and any(j % 2 == 1 for j in self._Jd):
            warnings.warn("accuracy may be compromised when L and J are both odd")

        self.alpha = [beatty_alpha(self._Jd[d], self._Kd[d], self._Nd[d]) for d in range(self.ndim)]
        self._sn_factors = [np.ascontiguousarray(
            deapod_factors(self._Nd[d], self._Kd[d], self._Jd[d], self.alpha[d], self.n_mid[d]), dtype=np.float64)
            for d in range(self.ndim)]
        self._lib = _lib.get_lib()
        self._plan = None
        self._create_plan()
        self._set_points(omega_np, omega_t)
        self.nargin1 = int(np.prod(self._Nd))
        self.nargout1 = self.M

    def _set_dtypes(self, precision):
        if precision == "single":
            self._cplx_dtype, self._real_dtype = np.dtype(np.complex64), np.dtype(np.float32)
        elif precision == "double":
            self._cplx_dtype, self._real_dtype = np.dtype(np.complex128), np.dtype(np.float64)
        else:
            raise ValueError("precision must be 'single', 'double' or 'auto'")
        self._precision = precision
        self._tcdt = torch.complex64 if precision == "single" else torch.complex128
        self._trdt = torch.float32 if precision == "single" else torch.float64

    def _create_plan(self):
        """Tables in the operator's precision and the device plan
        (mi355nufft_plan_create)."""
        tables = []
        for d in range(self.ndim):
            h = build_table(self._Nd[d], self._Jd[d], self._Kd[d], self._Ld, self._phasing)
            h = h.astype(self._cplx_dtype if self._phasing == "complex" else self._real_dtype)
            tables.append(np.ascontiguousarray(h))
        self._tables_host = tables
        self._h_dev_cache = None
        plan = ctypes.c_void_p()
        tab_ptrs = (ctypes.c_void_p * self.ndim)(*[t.ctypes.data for t in tables])
        sn_ptrs = (ctypes.POINTER(ctypes.c_double) * self.ndim)(
            *[s.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for s in self._sn_factors])
        phase_coef = _lib.dblarr([2 * np.pi / self._Kd[d] * self.n_mid[d] for d in range(self.ndim)])
        with torch.cuda.device(self.device):
            self._stream = torch.cuda.current_stream(self.device).cuda_stream
            _lib.check(self._lib.mi355nufft_plan_create(
                ctypes.byref(plan), self.ndim, _lib.i64arr(self._Nd), _lib.i64arr(self._Kd),
                _lib.i32arr(self._Jd), self._Ld, 0 if self._precision == "single" else 1,
                0 if self._phasing == "real" else 1, 1 if self.ortho else 0, float(self.adjoint_scalefactor),
                tab_ptrs, sn_ptrs, phase_coef if self._phasing == "real" else None, ctypes.c_void_p(self._stream)))
        self._plan = plan

    # ------------------------------------------------------------------
    # properties (names of the reference attributes)
    # ------------------------------------------------------------------
    Nd = property(lambda self: self._Nd)
    Kd = property(lambda self: self._Kd)
    Jd = property(lambda self: self._Jd)
    Ld = property(lambda self: self._Ld)
    mode = property(lambda self: self._mode)
    phasing = property(lambda self: self._phasing)
    # _nufft.py:510-513; the matrix itself lives on the device (rows for the
    # forward, columns for the adjoint), not as a host scipy matrix
    sparse_format = property(lambda self: "CSC" if self._sparse else None)

    @property
    def precision(self):
        """'single' or 'double'.  Setting it rebuilds the operator in the new
        precision (tables, plan, points; _nufft.py:520-545, where the setter
        re-casts omega, the phases, sn and the tables); binning, scatter and
        FFT choices return to their defaults."""
        return self._precision

    @precision.setter
    def precision(self, precision):
        if precision == "auto":
            precision = "single" if self._omega_user.dtype == torch.float32 else "double"
        if precision == self._precision:
            return
        old = (self._precision, self._plan)
        self._set_dtypes(precision)
        try:
            self._create_plan()
        except Exception:
            self._set_dtypes(old[0])
            self._plan = old[1]
            raise
        self._sync_stream()
        self._lib.mi355nufft_plan_destroy(old[1])
        frames = getattr(self, "_frame_offsets", None)
        omega_t, omega_np = self._omega_user, self._omega_np
        self._set_points(omega_np, omega_t, frames)
    n_shift = property(lambda self: self._n_shift)

    @property
    def omega(self):
        """Sampling points cast to the real dtype, (M, ndim) on the device."""
        return self._omega_rdt

    @property
    def h(self):
        """Interpolation tables (one per axis) as device tensors."""
        return [torch.from_numpy(t).to(self.device) for t in self._tables_host]

    @property
    def tm(self):
        """omega / (2 pi / Kd), (M, ndim), as in _nufft.py:338-342."""
        tm = torch.empty_like(self._omega_rdt)
        for d in range(self.ndim):
            gam = np.float64(2 * np.pi / self._Kd[d])
            g = torch.tensor(gam, dtype=torch.float64).to(self._trdt).item()
            tm[:, d] = (self._omega_rdt[:, d].double() / g).to(self._trdt)
        return tm

    @property
    def sn(self):
        """Deapodization factors sn[Nd] (float64 outer product cast to rdt)."""
        s = np.array([1.0])
        for f in self._sn_factors:
            s = np.outer(s.ravel(), f)
        return torch.from_numpy(s.reshape(self._Nd).astype(self._real_dtype)).to(self.device)

    @property
    def phase_before(self):
        """exp(i (2 pi n_mid / Kd) . k) evaluated in the real dtype (_nufft.py:703-715)."""
        if self._phasing != "real":
            return None
        rdt = self._real_dtype
        Kd, n_mid = self._Kd, self.n_mid
        phase = (2 * np.pi / Kd[0] * n_mid[0]) * np.arange(Kd[0], dtype=rdt)
        for d in range(1, self.ndim):
            tmp = (2 * np.pi / Kd[d] * n_mid[d]) * np.arange(Kd[d], dtype=rdt)
            phase = phase.reshape(phase.shape + (1,)) + tmp.reshape((1,) * d + (tmp.size,))
        return torch.from_numpy(np.exp(1j * phase).astype(self._cplx_dtype, copy=False)).to(self.device)

    @property
    def phase_after(self):
        """exp(i omega . (n_shift - n_mid)) from the uncast omega (_nufft.py:717-724)."""
        if self._phasing != "real":
            return None
        shift = np.asarray([s - m for s, m in zip(self._n_shift, self.n_mid)], dtype=self._real_dtype)
        # sparse mode recomputes it from the cast omega (_nufft.py:768-769 after :315)
        om = self._omega_rdt.cpu().numpy() if self._sparse else self._omega_host
        ph = np.exp(1j * np.dot(om, shift))
        return torch.from_numpy(ph.astype(self._cplx_dtype, copy=False)).to(self.device)

    # ------------------------------------------------------------------
    def _check_omega(self, omega):
        """_init_omega (_nufft.py:552-573)."""
        if isinstance(omega, torch.Tensor):
            omega_t = omega.detach().to(self.device)
            if omega_t.dim() == 1:
                omega_t = omega_t[:, None]
            omega_np = None
        else:
            omega_np = np.asarray(omega)
            if omega_np.ndim == 1:
                omega_np = omega_np[:, np.newaxis]
            if omega_np.dtype not in _SUPPORTED_REAL:
                raise ValueError("omega must be one of the following types: {}".format(
                    [np.float32, np.float64]))
            omega_t = torch.from_numpy(np.ascontiguousarray(omega_np)).to(self.device)
        if omega_t.dim() != 2 or omega_t.shape[1] != self.ndim:
            raise ValueError("number of cols must match NUFFT dimension")
        if omega_t.dtype not in (torch.float32, torch.float64):
            raise ValueError("omega must be one of the following types: {}".format([np.float32, np.float64]))
        return omega_np, omega_t

    @property
    def _omega_host(self):
        if self._omega_np is None:
            self._omega_np = self._omega_user.cpu().numpy()
        return self._omega_np

    def set_points(self, omega, frame_offsets=None):
        """Replace the sampling points (numpy array or device tensor (M, ndim)).

        Runs on the device (tm, per-sample phase, bin sort); the plan's
        tables, deapodization and FFT plans are kept.  ``frame_offsets``
        (length nframes + 1) splits the points into frames, each with its own
        grids (see ``NufftFrames``).
        """
        omega_np, omega_t = self._check_omega(omega)
        self._set_points(omega_np, omega_t, frame_offsets)

    def _set_points(self, omega_np, omega_t, frame_offsets=None):
        self.M = int(omega_t.shape[0])
        self.nargout1 = self.M
        self._omega_user = omega_t
        self._omega_np = omega_np
        self._omega_rdt = omega_t.to(self._trdt)
        self._points_version = getattr(self, "_points_version", 0) + 1
        if self._sparse and frame_offsets is not None:
            raise ValueError("sparse mode takes one frame")
        # sparse mode: phase_after from the cast omega (_nufft.py:768-769);
        # the complex kernel's phase_shift is folded into the matrix (:833-838)
        om = self._omega_rdt if self._sparse else omega_t
        if self._phasing == "real":
            shift = [float(np.asarray(s - m, dtype=self._real_dtype)) for s, m in zip(self._n_shift, self.n_mid)]
            phase_mode = 1
            dot_single = int(om.dtype == torch.float32 and self._precision == "single")
        elif any(s != 0 for s in self._n_shift) and not self._sparse:
            shift = list(self._n_shift)
            phase_mode = 2
            dot_single = 0
        else:
            shift = [0.0] * self.ndim
            phase_mode = 0
            dot_single = 0
        self._sync_stream()
        self._frame_offsets = frame_offsets
        if frame_offsets is None:
            offs, nframes = None, 1
        else:
            offs = _lib.i64arr(list(frame_offsets))
            nframes = len(frame_offsets) - 1
        _lib.check(self._lib.mi355nufft_set_points_frames(
            self._plan, ctypes.c_void_p(om.data_ptr()), 1 if om.dtype == torch.float64 else 0, self.M,
            om.stride(0), om.stride(1), _lib.dblarr(shift + [0.0] * (3 - self.ndim)), phase_mode, dot_single,
            offs, nframes))
        self.nframes = nframes
        if self._sparse:
            self._init_sparsemat()

    def _init_sparsemat(self):
        """Interpolation matrix of mode='sparse' (_init_sparsemat,
        _nufft.py:751-877).

        The per-axis factors ud[d] and grid indices kd[d] are computed here as
        in the reference, from omega cast to the real dtype (:772-802 with
        _nufft_coef / _nufft_offset, _utils.py:20-86); the device forms the
        entries prod_d ud[d] in float64 / complex128, conjugated for the
        complex kernel and times exp(i omega . n_shift) (:815-838), rounded to
        the operator's precision (``mi355nufft_set_sparse``).
        """
        om = self._omega_rdt.cpu().numpy()
        cplx = self._phasing == "complex"
        coefs, kidx = [], []
        for d in range(self.ndim):
            N, J, K = self._Nd[d], self._Jd[d], self._Kd[d]
            o = om[:, d]
            gam = 2 * np.pi / K
            koff = np.floor(o / gam - J / 2.0)                          # _utils.py:40-41
            dk = o / gam - koff                                         # _utils.py:74-75
            arg = -np.arange(1, J + 1)[:, None] + dk[None, :]           # _utils.py:78, (J, M)
            c = kaiser_bessel(arg, J, self.alpha[d])
            kd = np.mod(np.arange(1, J + 1)[:, None] + koff[None, :], K)   # :791
            if cplx:
                phase_scale = 1j * gam * (N - 1) / 2.0                  # :793-796
                ud = np.exp(phase_scale * arg) * c
            else:
                ud = c
            ud = np.ascontiguousarray(ud, dtype=np.complex128 if cplx else np.float64)
            coefs.append(torch.from_numpy(ud).to(self.device))
            kidx.append(torch.from_numpy(np.ascontiguousarray(kd.astype(np.int32))).to(self.device))
        phase = None
        if cplx and any(s != 0 for s in self._n_shift):                 # :833-838
            phase = torch.from_numpy(np.exp(1j * np.dot(om, np.asarray(self._n_shift)))).to(self.device)
        cp = (ctypes.c_void_p * self.ndim)(*[t.data_ptr() for t in coefs])
        kp = (ctypes.c_void_p * self.ndim)(*[t.data_ptr() for t in kidx])
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_set_sparse(self._plan, cp, kp,
                                                   None if phase is None else ctypes.c_void_p(phase.data_ptr())))

    _METHODS = {"auto": 0, "gm": 1, "tiled": 2, "tiled_atomic": 3, "tiled_two_phase": 4}

    def set_binning(self, bin_size=None, max_points=0, method="auto", threads=0):
        """Bin geometry and kernel family of the interpolation stages.

        ``bin_size``: per-axis bin size in grid points (None: defaults);
        ``max_points``: points per subproblem (0: default); ``method``:
        'auto' (tiled when the tile fits in LDS), 'gm' (global-memory kernels)
        'tiled', 'tiled_atomic' (scatter flushed with global atomics) or
        'tiled_two_phase' (scatter through tile slots and a deterministic
        per-grid-point reduction).  ``threads``: workgroup size of the tiled
        kernels (64, 128, 256; 0: default).  The current points are re-sorted.
        """
        bs = None
        if bin_size is not None:
            bs = _lib.i32arr(_as_tuple(bin_size, type=int, n=self.ndim))
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_set_tiling(self._plan, bs, int(max_points), self._METHODS[method],
                                                   int(threads)))

    def plan_info(self):
        """dict with the binning state of the plan (see mi355nufft_plan_info)."""
        buf = (ctypes.c_int64 * 19)()
        _lib.check(self._lib.mi355nufft_plan_info(self._plan, buf, 19))
        v = list(buf)
        return dict(ndim=v[0], M=v[1], bin_size=tuple(v[2:5]), max_points=v[5], nbins=v[6], tile=v[7],
                    tiled=bool(v[8]), subproblem_slots=v[9], method=v[10], lds_limit=v[11], nframes=v[12],
                    two_phase=bool(v[13]), threads=v[14], fused_fft=bool(v[15]),
                    sparse_nnz=v[16] if v[16] >= 0 else None, sparse_pieces=v[17], fixed_point_scatter=bool(v[18]))

    def set_scatter(self, accumulate="auto"):
        """Accumulation of the scatter's LDS tile: "auto" (32-bit fixed point
        for 2-D and 3-D complex64, float64 otherwise), "float64", or "fixed32"
        (32-bit fixed point for any complex64 plan; see
        mi355nufft_set_scatter)."""
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_set_scatter(self._plan, {"auto": 0, "float64": 1, "fixed32": 2}[accumulate]))

    def set_fft(self, method="auto"):
        """FFT of the transforms: "auto" (pruned passes fused with the
        deapodization for power-of-two grids, 16 <= K <= 2048, complex128
        K <= 1024) or "rocfft"."""
        m = {"auto": 0, "rocfft": 1}[method]
        _lib.check(self._lib.mi355nufft_set_fft(self._plan, m))

    def set_kernel_timing(self, enable=True, count_atomics=False):
        """Bracket every tiled gather / scatter launch with HIP events on the
        plan's stream (skipped while the stream is being captured);
        ``count_atomics`` also counts the scatter flush's global atomics
        (``read_flush_atomics``)."""
        v = (3 if count_atomics else 1) if enable else 0
        _lib.check(self._lib.mi355nufft_set_timing(self._plan, v))

    def read_kernel_timing(self, which):
        """(launches, mean_ms) of the gather ("interp") or scatter ("spread")
        launches timed since the last read; waits for them to finish."""
        w = {"interp": 0, "spread": 1}[which]
        n = ctypes.c_int64(0)
        ms = ctypes.c_double(0.0)
        _lib.check(self._lib.mi355nufft_read_timing(self._plan, w, ctypes.byref(n), ctypes.byref(ms)))
        return n.value, (ms.value / n.value if n.value else float("nan"))

    def read_flush_atomics(self):
        """Global float atomic lane operations made by the tile flushes of the
        scatter launches timed since the last read (then reset)."""
        v = ctypes.c_uint64(0)
        _lib.check(self._lib.mi355nufft_read_flush_atomics(self._plan, ctypes.byref(v)))
        return v.value

    def _h_dev(self):
        if getattr(self, "_h_dev_cache", None) is None:
            self._h_dev_cache = [torch.from_numpy(t).to(self.device) for t in self._tables_host]
        return self._h_dev_cache

    def _tm_dev(self):
        """tm as (ndim, M) contiguous, the layout of the raw table1 kernels."""
        key = self._points_version
        if getattr(self, "_tm_dev_key", None) != key:
            self._tm_dev_cache = self.tm.t().contiguous()
            self._tm_dev_key = key
        return self._tm_dev_cache

    def _sync_stream(self):
        s = torch.cuda.current_stream(self.device).cuda_stream
        if s != self._stream:
            _lib.check(self._lib.mi355nufft_set_stream(self._plan, ctypes.c_void_p(s)))
            self._stream = s

    def __del__(self):
        plan = getattr(self, "_plan", None)
        lib = getattr(self, "_lib", None)
        if plan is not None and lib is not None and plan.value:
            try:
                lib.mi355nufft_plan_destroy(plan)
            except Exception:
                pass
            self._plan = None

    # ------------------------------------------------------------------
    def _swap_reps(self, x, narg):
        """_nufft.py:415-418."""
        if _numel(x) != narg:
            x = x.permute(*(tuple(range(1, x.dim())) + (0,))) if isinstance(x, torch.Tensor) else \
                x.transpose(tuple(range(1, x.ndim)) + (0,))
        return x

    def _unswap_reps(self, x, narg):
        """_nufft.py:420-423."""
        if _numel(x) != narg:
            x = x.permute(*((x.dim() - 1,) + tuple(range(x.dim() - 1)))) if isinstance(x, torch.Tensor) else \
                x.transpose((x.ndim - 1,) + tuple(range(x.ndim - 1)))
        return x

    def fft(self, x):
        """Forward NUFFT (uniform spatial -> non-uniform frequency), _nufft.py:425-452."""
        if self.order == "C":
            x = self._swap_reps(x, self.nargin1)
        k = nufft_forward(self, x)
        if self.order == "C":
            k = self._unswap_reps(k, self.nargout1)
        return k

    def adj(self, k):
        """Adjoint NUFFT (non-uniform frequency -> uniform spatial), _nufft.py:454-482."""
        if self.order == "C":
            k = self._swap_reps(k, self.nargout1)
        x = nufft_adj(self, k)
        if self.order == "C":
            x = self._unswap_reps(x, self.nargin1)
        return x

    # ------------------------------------------------------------------
    # device-tensor entry points used by autograd / batching
    # ------------------------------------------------------------------
    def _forward_dev(self, x, nreps):
        """x: device tensor, logical shape Nd + (nreps,) (any strides); returns (nreps, M)."""
        y = torch.empty((nreps, self.M), dtype=self._tcdt, device=self.device)
        grid = torch.empty((self.nframes * nreps, int(np.prod(self._Kd))), dtype=self._tcdt, device=self.device)
        xs = list(x.stride())
        self._sync_stream()
        _lib.check(self._lib.mi355nufft_forward(
            self._plan, ctypes.c_void_p(x.data_ptr()), _lib.i64arr(xs), ctypes.c_void_p(y.data_ptr()),
            nreps, ctypes.c_void_p(grid.data_ptr())))
        return y

    def _adjoint_dev(self, y, nreps, out=None):
        """y: device tensor (nreps, M) contiguous; returns Nd + (nreps,) Fortran-ordered."""
        if out is None:
            out = torch.empty((nreps,) + tuple(reversed(self._Nd)), dtype=self._tcdt, device=self.device)
            out = out.permute(*(tuple(range(self.ndim, 0, -1)) + (0,)))
        self._sync_stream()
        # plan-owned grid: left zero by each adjoint's last FFT stage
        _lib.check(self._lib.mi355nufft_adjoint(
            self._plan, ctypes.c_void_p(y.data_ptr()), ctypes.c_void_p(out.data_ptr()),
            _lib.i64arr(list(out.stride())), nreps, None))
        return out


def _numel(x):
    return x.numel() if isinstance(x, torch.Tensor) else np.size(x)


_TORCH_DTYPE = {np.dtype(np.complex64): torch.complex64, np.dtype(np.complex128): torch.complex128,
                np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64}


def _upload_host(x, device):
    """numpy array -> device tensor of the same shape, dtype and memory order
    (floating or complex dtypes; others as float64), asynchronous.

    The array is copied into page-locked memory from torch's caching host
    allocator (a cached block, not a new pinned allocation per call), as one
    flat memcpy in the array's own order (C or Fortran), and that flat
    buffer goes to the device in one DMA; the device tensor then takes the
    array's strides.  The allocator keeps the block until the copy has
    completed."""
    x = np.asarray(x)
    if x.dtype not in _TORCH_DTYPE:
        x = x.astype(np.float64)
    if not (x.flags.c_contiguous or x.flags.f_contiguous):
        x = np.ascontiguousarray(x)
    tdt = _TORCH_DTYPE[x.dtype]
    pinned = torch.empty(max(x.size, 1), dtype=tdt, pin_memory=True)[:x.size]
    order = "C" if x.flags.c_contiguous else "F"
    np.copyto(pinned.numpy(), x.reshape(-1, order=order), casting="no")
    flat = pinned.to(device, non_blocking=True)
    if order == "C":
        return flat.reshape(x.shape)
    return flat.reshape(x.shape[::-1]).permute(*reversed(range(x.ndim)))


def _to_device(obj, x):
    """Returns (tensor on device with the operator's complex dtype, was_numpy).

    Host arrays go through the device's persistent pinned staging buffer,
    so the transfer runs at pinned-memory rate without a pinned allocation
    per call."""
    was_numpy = not isinstance(x, torch.Tensor)
    if was_numpy:
        x = np.asarray(x)
        if x.dtype == np.complex128 and obj._precision == "single":
            x = x.astype(np.complex64)
        t = _upload_host(x, obj.device)
    else:
        t = x.to(obj.device)
    if not t.is_complex():
        t = t.to(obj._trdt)
        t = torch.complex(t, torch.zeros_like(t))
    if t.dtype != obj._tcdt:
        t = t.to(obj._tcdt)
    return t, was_numpy


class _OutStaging(object):
    """Persistent page-locked landing buffer of one device for results going
    back to numpy callers, grown on demand.  A result arrives in it by one
    DMA at pinned rate and is then copied into an ordinary numpy array, so
    callers never hold page-locked memory."""

    def __init__(self):
        self.lock = threading.Lock()
        self.buf = None

    def download(self, t):
        nbytes = t.numel() * t.element_size()
        with self.lock:
            if self.buf is None or self.buf.numel() < nbytes:
                self.buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, pin_memory=True)
            flat = self.buf[:nbytes].view(t.dtype)
            if t.is_contiguous():
                staged = flat.reshape(t.shape)
            else:  # Fortran order
                staged = flat.reshape(tuple(reversed(t.shape))).permute(*reversed(range(t.dim())))
            staged.copy_(t, non_blocking=True)
            torch.cuda.current_stream(t.device).synchronize()
            return np.array(staged.numpy(), copy=True, order="K")


_OUT_STAGING = {}
_OUT_STAGING_LOCK = threading.Lock()


def _to_output(t, was_numpy):
    """Device results for numpy callers: a new ordinary numpy array
    (synchronous).  Dense results (C or Fortran order) go through the
    device's persistent pinned landing buffer and keep their memory order;
    other layouts take a plain device-to-host copy."""
    if not was_numpy:
        return t
    if t.is_contiguous() or t.permute(*reversed(range(t.dim()))).is_contiguous():
        key = t.device.index if t.device.index is not None else torch.cuda.current_device()
        with _OUT_STAGING_LOCK:
            st = _OUT_STAGING.get(key)
            if st is None:
                st = _OUT_STAGING[key] = _OutStaging()
        return st.download(t)
    return t.cpu().numpy()


def _raw_table1(obj, fwd, data, nreps):
    """Raw table1 gather (fwd) or scatter on the plan's tables and points."""
    h = obj._h_dev()
    hp = (ctypes.c_void_p * obj.ndim)(*[t.data_ptr() for t in h])
    tm = obj._tm_dev()
    obj._sync_stream()
    stream = ctypes.c_void_p(obj._stream)
    prec = 0 if obj._precision == "single" else 1
    ck = 1 if obj._phasing == "complex" else 0
    geom = (obj.ndim, _lib.i64arr(obj.Kd), _lib.i32arr(obj.Jd), obj.Ld, prec, ck)
    if fwd:
        out = torch.empty((nreps, obj.M), dtype=obj._tcdt, device=obj.device)
        _lib.check(obj._lib.mi355nufft_table1_forward(*geom, ctypes.c_void_p(data.data_ptr()), hp,
                                                      ctypes.c_void_p(tm.data_ptr()), obj.M,
                                                      ctypes.c_void_p(out.data_ptr()), nreps, stream))
    else:
        out = torch.empty((nreps, int(np.prod(obj.Kd))), dtype=obj._tcdt, device=obj.device)
        _lib.check(obj._lib.mi355nufft_table1_adjoint(*geom, ctypes.c_void_p(data.data_ptr()), hp,
                                                      ctypes.c_void_p(tm.data_ptr()), obj.M,
                                                      ctypes.c_void_p(out.data_ptr()), nreps, stream))
    return out


def _sparse_apply(obj, adjoint, data, nreps):
    """Bare sparse-matrix product: P data (forward) or P^H data (adjoint)."""
    if adjoint:
        out = torch.empty((nreps, int(np.prod(obj.Kd))), dtype=obj._tcdt, device=obj.device)
    else:
        out = torch.empty((nreps, obj.M), dtype=obj._tcdt, device=obj.device)
    obj._sync_stream()
    _lib.check(obj._lib.mi355nufft_sparse_apply(obj._plan, 1 if adjoint else 0, ctypes.c_void_p(data.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()), nreps))
    return out


def _phase_shift_dev(obj):
    """Complex-kernel phase_shift exp(i omega_rdt . n_shift), complex128 (_nufft.py:898-901)."""
    if obj._phasing != "complex" or not any(s != 0 for s in obj.n_shift):
        return None
    sh = torch.tensor(obj.n_shift, dtype=torch.float64, device=obj.device)
    return torch.exp(1j * (obj._omega_rdt.double() @ sh))


def _check_xp(xp):
    """The reference's array-module argument: numpy (or None) selects numpy
    results for numpy inputs, torch keeps device tensors; the transform runs
    on the GPU either way."""
    if xp is not None and xp is not np and xp is not torch:
        raise ValueError("xp must be None, numpy or torch")


def nufft_forward(obj, x, copy_x=True, grid_only=False, xp=None):
    """Forward table1 NUFFT (mrrt/nufft/_nufft.py:1275-1397), same signature.

    ``copy_x``: the input is never modified (it is read by the device), so
    no copy is needed either way.  ``grid_only``: ``x`` is the oversampled
    grid (prod(Kd), reps); only the table interpolation is applied (no FFT,
    phase_before, ortho or phase_after; the complex kernel's phase_shift is
    kept, :1086-1095) and the (M, reps) result is returned (:1302-1304,
    :1376-1388).  ``xp``: see ``_check_xp``.
    """
    _check_xp(xp)
    if grid_only:
        t, was_numpy = _to_device(obj, x)
        nk = int(np.prod(obj.Kd))
        if t.numel() % nk != 0:
            raise ValueError("Input signal has the wrong size.")
        nreps = t.numel() // nk
        g = _f_reshape(t, (nk, nreps)).t().contiguous()
        if obj._sparse:                       # x = obj.p * xk (:1384), phase_shift inside p
            return _to_output(_sparse_apply(obj, False, g, nreps).t(), was_numpy)
        y = _raw_table1(obj, True, g, nreps)
        ph = _phase_shift_dev(obj)
        if ph is not None:
            y = (y.to(torch.complex128) * ph[None, :]).to(obj._tcdt)
        return _to_output(y.t(), was_numpy)
    t, was_numpy = _to_device(obj, x)
    nn = obj.nargin1
    if t.numel() % nn != 0:
        raise ValueError("Input signal has the wrong size.")
    nreps = t.numel() // nn
    xs = _f_reshape(t, tuple(obj.Nd) + (nreps,))
    y = obj._forward_dev(xs, nreps)           # (nreps, M) = Fortran (M, nreps)
    out = y.t()
    if nreps == 1:
        out = out[:, 0]
    return _to_output(out, was_numpy)


def nufft_adj(obj, xk, copy=True, return_psf=False, grid_only=False, xp=None):
    """Adjoint table1 NUFFT (mrrt/nufft/_nufft.py:1459-1578), same signature.

    ``copy``: the input is never modified, so no copy is needed either way.
    ``grid_only``: return the raw gridded samples (prod(Kd), reps) before
    conj(phase_before) and the FFT (:1507-1509); it is checked first, as in
    the reference, so with ``return_psf`` as well the flat grid comes back
    (without conj(phase_after)).  ``return_psf`` alone: skip conj(phase_after)
    and return the first repetition's grid with shape Kd (:1495, :1517-1518),
    e.g. for a Toeplitz point-spread function.  ``xp``: see ``_check_xp``.
    """
    _check_xp(xp)
    t, was_numpy = _to_device(obj, xk)
    if t.numel() % obj.M != 0:
        raise ValueError("invalid size")
    nreps = t.numel() // obj.M
    y = _f_reshape(t, (obj.M, nreps)).t().contiguous()   # (nreps, M)
    if grid_only or return_psf:
        if obj._phasing == "real" and not return_psf:
            y = y * torch.conj(obj.phase_after)[None, :]
        if obj._sparse:                       # xk = obj.p.H * xk (:1505), phase_shift inside p
            g = _sparse_apply(obj, True, y.contiguous(), nreps)
        else:
            ph = _phase_shift_dev(obj)
            if ph is not None:
                y = (y.to(torch.complex128) * torch.conj(ph)[None, :]).to(obj._tcdt)
            g = _raw_table1(obj, False, y.contiguous(), nreps)
        if grid_only:
            return _to_output(g.t(), was_numpy)
        g0 = g[0].reshape(tuple(reversed(obj.Kd))).permute(*reversed(range(obj.ndim)))
        return _to_output(g0, was_numpy)
    out = obj._adjoint_dev(y, nreps)
    if nreps == 1:
        out = out[..., 0]
    return _to_output(out, was_numpy)
