"""numpy/scipy restatement of the reference table1 NUFFT (the oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Every function cites the reference file:line it restates (paths relative to
the reference tree, package ``mrrt/nufft``).  The float32 quirks of the
reference are kept on purpose, because parity is measured against the
reference's own output:

* the interpolation table is built through a float32 sparse 1-D NUFFT
  (``_nufft.py:1198-1243`` with the default ``precision="single"`` of
  ``_nufft.py:222``), so its values are float32 and its sample positions are
  float32-rounded, even for a double-precision operator;
* ``phase_before`` is evaluated in the operator's real dtype
  (``_nufft.py:703-715``);
* ``tm = omega_rdt / gamma`` is evaluated after the cast of omega to the real
  dtype (``_nufft.py:315`` then ``:338-342``);
* ``phase_after`` is evaluated from the *uncast* omega (``_nufft.py:313``).

Pinning (what checks each function against something other than itself):

=====================  ====================================================
kaiser_bessel          closed form I0 KB; support/zero checks of
                       test_kernels.py:56-58; brute-force scipy.special.iv
kaiser_bessel_ft       closed form J*sinh(w)/(I0(a) w); FFT of sampled KB
beatty_alpha           closed form (Beatty 2005 eq. 5), alpha(6, 2) value
nufft_offset           KAT of test_utils.py:18-22
table_make1            slow == fast == ratio (test_nufft.py:391-409);
                       float64 closed-form KB at the float32 positions;
                       independent direct builder in the product
                       (bit-exact); the float32 *position* rounding itself
                       is pinned only by reading _nufft.py (no import of
                       the reference is possible: mrrt.utils is absent)
interp_forward/adj     reference C kernels (oracle/_ref) on golden vectors;
                       adjointness; brute-force tap sum on tiny inputs
dtft / dtft_adj        FFT at uniform frequencies (test_dtft.py) atol 1e-12
NufftOracle.fft/adj    composition of the above + numpy.fft; vs dtft at the
                       reference's own tolerances (test_nufft.py:60-324)
sparse_matrix          every entry vs a brute-force sum over the whole grid
                       (1-D/2-D/3-D, real and complex kernel, n_shift); the
                       sparse-mode operator vs dtft.  Sparse mode's
                       phase_after from the cast omega (_nufft.py:768-769) is
                       pinned only by reading the source
=====================  ====================================================
"""

from math import sqrt

import numpy as np
import scipy.sparse
from scipy.special import i0, jv

__all__ = [
    "beatty_alpha",
    "kaiser_bessel",
    "kaiser_bessel_ft",
    "nufft_offset",
    "nufft_coef",
    "table_make1",
    "calc_scaling_1d",
    "phase_before",
    "phase_after",
    "compute_tm",
    "interp_forward",
    "interp_adjoint",
    "dtft",
    "dtft_adj",
    "NufftOracle",
]


# --------------------------------------------------------------------------
# kernel math (setup time)
# --------------------------------------------------------------------------
def beatty_alpha(J, K, N):
    """Beatty KB shape parameter, ``_kernels.py:151-154`` (m = 0, ``:160``)."""
    k_n = K / N
    return np.pi * sqrt(J ** 2 / k_n ** 2 * (k_n - 0.5) ** 2 - 0.8)


def kaiser_bessel(x, J, alpha):
    """Order-0 KB window on [-J/2, J/2], ``_kaiser_bessel.py:78-149``.

    Follows the m == 0 branch (``:137-147``) operation for operation:
    ``tmp = 2 x / J; tmp *= tmp; f = sqrt(1 - tmp); i0(alpha f) / i0(alpha)``.
    """
    x = np.asarray(x)
    ii = (2 * np.abs(x) < J).nonzero()
    tmp = 2 * x[ii] / J
    tmp *= tmp
    f = np.sqrt(1 - tmp)
    denom = i0(alpha)
    kb = np.zeros_like(x)
    kb[ii] = i0(alpha * f) / float(denom)
    return kb.real


def kaiser_bessel_ft(u, J, alpha, d=1):
    """FT of the order-0 KB window, ``_kaiser_bessel.py:153-226`` (m = 0).

    Restated through the same complex-Bessel route as the reference
    (``scimath.sqrt`` then ``jv(d/2, z) / z**(d/2)``).
    """
    u = np.asarray(u, dtype=np.float64)
    tmp = (np.pi * J) * u
    tmp *= tmp
    tmp -= alpha * alpha
    z = np.lib.scimath.sqrt(tmp)
    nu = d / 2.0
    const1 = (2 * np.pi) ** (d / 2.0) * (J / 2.0) ** d
    const1 /= i0(alpha)
    y = const1 * jv(nu, z)
    y /= z ** nu
    return np.real(y)


def nufft_offset(om, J, K):
    """``_utils.py:20-43``: ``floor(om / gamma - J / 2)``, dtype of ``om``."""
    om = np.asanyarray(om)
    gam = 2 * np.pi / K
    return np.floor(om / gam - J / 2.0)


def nufft_coef(om, J, K, alpha):
    """``_utils.py:47-86`` with the Beatty KB kernel: returns (coef, arg)."""
    om = np.atleast_1d(np.squeeze(om))
    gam = 2 * np.pi / K
    dk = om / gam - nufft_offset(om, J, K)
    arg = -np.arange(1, J + 1)[:, None] + dk[None, :]
    return kaiser_bessel(arg, J, alpha), arg


def _sparse_1d_matrix(omega, N, J, K, phasing):
    """Sparse interpolation matrix of a 1-D float32 NufftBase.

    Restates ``NufftBase.__init__`` (mode='sparse', default precision
    'single', ``_nufft.py:216-322``) followed by ``_init_sparsemat``
    (``_nufft.py:751-877``) for ndim == 1 and n_shift == 0.
    """
    alpha = beatty_alpha(J, K, N)
    om32 = np.asarray(omega, dtype=np.float64).astype(np.float32)  # :315
    c, arg = nufft_coef(om32, J, K, alpha)                           # :783
    koff = nufft_offset(om32, J, K)                                  # :788
    kd = np.mod(np.arange(1, J + 1)[:, None] + koff[None, :], K)     # :791
    if phasing == "complex":                                         # :793-796
        gam = 2 * np.pi / K
        phase_scale = 1j * gam * (N - 1) / 2.0
        phase = np.exp(phase_scale * arg)
    else:
        phase = 1.0
    uu = phase * c                                                   # :802
    if np.iscomplexobj(uu):                                          # :830-831
        uu = uu.conj()
    sparse_dtype = np.complex64 if phasing == "complex" else np.float32
    M = om32.shape[0]
    mm = np.tile(np.arange(M), (J, 1))                               # :852
    p = scipy.sparse.coo_matrix(
        (uu.ravel(order="F"),
         (mm.ravel(order="F"), kd.astype(np.int64).ravel(order="F"))),
        shape=(M, K),
        dtype=sparse_dtype,
    )
    return p.tocsc()


def sparse_matrix(omega_rdt, Nd, Jd, Kd, alpha, phasing, n_shift, rdt):
    """Interpolation matrix of ``mode='sparse'``: ``_init_sparsemat``
    (``_nufft.py:751-877``) for any ndim, as a scipy CSC matrix (M, prod Kd).

    ``omega_rdt`` is omega already cast to the real dtype (``:315`` runs
    before ``:322``).  Products in float64 / complex128, conjugated for the
    complex kernel, times exp(i omega . n_shift), then cast to the matrix
    dtype (real dtype, or complex dtype for the complex kernel).
    """
    omega = np.asarray(omega_rdt)
    if omega.ndim == 1:
        omega = omega[:, None]
    ndim = omega.shape[1]
    M = omega.shape[0]
    ud, kd = [], []
    for d in range(ndim):                                            # :775-802
        N, J, K = Nd[d], Jd[d], Kd[d]
        c, arg = nufft_coef(omega[:, d], J, K, alpha[d])
        koff = nufft_offset(omega[:, d], J, K)
        kd.append(np.mod(np.arange(1, J + 1)[:, None] + koff[None, :], K))
        if phasing == "complex":
            gam = 2 * np.pi / K
            phase_scale = 1j * gam * (N - 1) / 2.0
            phase = np.exp(phase_scale * arg)
        else:
            phase = 1.0
        ud.append(phase * c)
    kk, uu = kd[0], ud[0]                                            # :812-823
    for d in range(1, ndim):
        Jprod = int(np.prod(Jd[: d + 1]))
        tmp = kd[d] * int(np.prod(Kd[:d]))
        kk = (kk[:, None, :] + tmp[None, :, :]).reshape((Jprod, M), order="F")
        uu = (uu[:, None, :] * ud[d][None, :, :]).reshape((Jprod, M), order="F")
    if np.iscomplexobj(uu):                                          # :830-831
        uu = uu.conj()
    if phasing == "complex":                                         # :833-840
        if any(s != 0 for s in n_shift):
            phase = np.exp(1j * np.dot(omega, np.asarray(n_shift))).reshape((1, -1), order="F")
            uu = uu * phase
        sparse_dtype = np.complex64 if np.dtype(rdt) == np.float32 else np.complex128
    else:
        sparse_dtype = rdt
    mm = np.tile(np.arange(M), (int(np.prod(Jd)), 1))               # :852
    p = scipy.sparse.coo_matrix(
        (uu.ravel(order="F"), (mm.ravel(order="F"), kk.astype(np.int64).ravel(order="F"))),
        shape=(M, int(np.prod(Kd))),
        dtype=sparse_dtype,
    )
    return p.tocsc()


def table_make1(how, N, J, K, L, phasing="real"):
    """Interpolation table ``h[J*L+1]``, ``_nufft.py:1195-1271``.

    Returns ``(h, t0)`` with ``h`` float32 (real phasing) or complex64.
    """
    t0 = np.arange(-J * L / 2.0, J * L / 2.0 + 1) / L
    if how == "slow":                                                # :1217-1220
        omega0 = t0 * 2 * np.pi / K
        p = _sparse_1d_matrix(omega0, N, J, K, phasing)
        h = np.asarray(p[:, 0].todense()).ravel()
    elif how == "fast":                                              # :1223-1243
        if N % 2 == 0:
            t1 = J / 2.0 - 1 + np.arange(L) / L
        else:
            t1 = J / 2.0 - 1 + np.arange(1, L + 1) / L
        omega1 = t1 * 2 * np.pi / K
        p = _sparse_1d_matrix(omega1, N, J, K, phasing)
        h = np.asarray(p[:, np.arange(J - 1, -1, -1)].todense()).ravel(order="F")
        if N % 2 == 0:
            h = np.concatenate((h, np.atleast_1d(h[0])), axis=0)
        else:
            h = np.concatenate((np.atleast_1d(h[-1]), h), axis=0)
    elif how == "ratio":                                             # :1247-1268
        Nfake = J
        Kfake = Nfake * K / N
        t1 = J / 2.0 - 1 + np.arange(L) / L
        omega1 = t1 * 2 * np.pi / Kfake
        # the nested NufftBase truncates Kd to int (_utils.py:130-137)
        p = _sparse_1d_matrix(omega1, Nfake, J, int(Kfake), phasing)
        h = np.asarray(p[:, np.arange(J - 1, -1, -1)].todense()).ravel(order="F")
        h = np.concatenate((h, np.asarray([h[0]])), axis=0)
        if phasing == "complex":
            h = h * np.exp(1j * np.pi * t0 * (1 / K - 1 / Kfake))
    else:
        raise ValueError("Bad Type: {}".format(how))
    return h, t0


def calc_scaling_1d(N, K, J, alpha, n_mid):
    """One axis of ``_calc_scaling`` (``_nufft.py:727-748``), float64."""
    start = -n_mid
    nc = np.arange(start, start + N)
    tmp = 1 / kaiser_bessel_ft(nc / K, J, alpha, 1)
    return np.real(tmp).conj()


# --------------------------------------------------------------------------
# phases and normalised coordinates (setup time)
# --------------------------------------------------------------------------
def phase_before(Kd, n_mid, rdt):
    """``_nufft.py:703-715``: evaluated in the real dtype ``rdt``."""
    rdt = np.dtype(rdt)
    cdt = np.result_type(rdt, np.complex64)
    phase = (2 * np.pi / Kd[0] * n_mid[0]) * np.arange(Kd[0], dtype=rdt)
    for d in range(1, len(Kd)):
        tmp = (2 * np.pi / Kd[d] * n_mid[d]) * np.arange(Kd[d], dtype=rdt)
        phase = phase.reshape(phase.shape + (1,)) + tmp.reshape((1,) * d + (tmp.size,))
    return np.exp(1j * phase).astype(cdt, copy=False)


def phase_after(omega, n_mid, n_shift, rdt):
    """``_nufft.py:717-724``; ``omega`` is the *uncast* user array."""
    rdt = np.dtype(rdt)
    cdt = np.result_type(rdt, np.complex64)
    shift_vec = [(s - m) for s, m in zip(n_shift, n_mid)]
    phase = np.exp(1j * np.dot(omega, np.asarray(shift_vec, dtype=rdt)))
    return phase.astype(cdt, copy=False)


def compute_tm(omega_rdt, Kd):
    """``_nufft.py:338-342`` on the already-cast omega."""
    tm = np.zeros_like(omega_rdt)
    for d in range(len(Kd)):
        gam = 2 * np.pi / Kd[d]
        tm[:, d] = omega_rdt[:, d] / gam
    return tm


# --------------------------------------------------------------------------
# table1 interpolation (the hot loops)
# --------------------------------------------------------------------------
def _axis_taps(t, J, L, K, h, cplx_kernel):
    """Per-axis tap indices and coefficients of the C kernels.

    ``nufft_table.template.c:577-601`` (real kernel) and ``:662-688``
    (complex kernel): ``k = 1 + floor(t - J/2.)`` in double;
    ``p = (t - k) * L``, ``n = floor(p)``, ``alf = p - n`` in the real dtype;
    ``coef = (1 - alf) h[n + nc] + alf h[n + 1 + nc]``, ``nc = floor(J L / 2)``.
    Both wrap forms of the C code (float ``floor(k / K)`` for the real
    kernel, double ``mymod`` for the complex kernel) equal the exact modulo.
    Returns ``kmod`` (M, J) int64 and ``coef`` (M, J).

    Rounding edge: when ``t - k`` rounds to exactly ``J/2`` (t just below an
    integer with |t| < 2), ``p == J*L/2`` and the C code reads ``h[J*L+1]``,
    one element past the table, with weight ``alf == 0``.  The restatement
    reads an appended zero there, i.e. the value the C code produces whenever
    that out-of-range word is finite.  For odd ``J*L`` (warned as inaccurate
    at _nufft.py:330-335) a cell below the table start is clamped to it.
    """
    T = t.dtype.type
    k0 = 1 + np.floor(t.astype(np.float64) - J / 2.0)
    ks = k0[:, None].astype(np.int64) + np.arange(J)[None, :]
    p = (t[:, None] - ks.astype(t.dtype)) * T(L)
    nc = (J * L) // 2
    n = np.maximum(np.floor(p).astype(np.int64), -nc)
    alf = p - n.astype(t.dtype)
    h = np.concatenate((h, np.zeros(1, h.dtype)))
    idx = n + nc
    if cplx_kernel:
        coef = (1 - alf) * h.real[idx] + alf * h.real[idx + 1]
        coef = coef + 1j * ((1 - alf) * h.imag[idx] + alf * h.imag[idx + 1])
        coef = coef.astype(np.result_type(t.dtype, np.complex64))
    else:
        coef = (1 - alf) * h[idx] + alf * h[idx + 1]
    return np.mod(ks, K), coef


def interp_forward(ck, Kd, hs, Jd, L, tm, cplx_kernel=False):
    """Forward table1 gather, ``nufft_table.template.c:43-141, 538-703, 710-921``.

    ``ck``: (prod(Kd), nreps) complex, Fortran grid index
    ``k1 + K1 k2 + K1 K2 k3``; ``tm``: (M, ndim) real.  Returns (M, nreps).
    Sums run in the same nesting as the C code (innermost axis 0).
    """
    ndim = len(Kd)
    ck = np.asarray(ck)
    if ck.ndim == 1:
        ck = ck[:, None]
    M = tm.shape[0]
    taps = [_axis_taps(tm[:, d], Jd[d], L, Kd[d], hs[d], cplx_kernel)
            for d in range(ndim)]
    strides = [int(np.prod(Kd[:d])) for d in range(ndim)]
    out = np.zeros((M, ck.shape[1]), dtype=ck.dtype)
    for r in range(ck.shape[1]):
        g = ck[:, r]
        if ndim == 1:
            (k1, c1), = taps
            s = np.zeros(M, ck.dtype)
            for j1 in range(Jd[0]):
                s += c1[:, j1] * g[k1[:, j1]]
            out[:, r] = s
        elif ndim == 2:
            (k1, c1), (k2, c2) = taps
            s2 = np.zeros(M, ck.dtype)
            for j2 in range(Jd[1]):
                s1 = np.zeros(M, ck.dtype)
                base = k2[:, j2] * strides[1]
                for j1 in range(Jd[0]):
                    s1 += c1[:, j1] * g[base + k1[:, j1]]
                s2 += c2[:, j2] * s1
            out[:, r] = s2
        else:
            (k1, c1), (k2, c2), (k3, c3) = taps
            s3 = np.zeros(M, ck.dtype)
            for j3 in range(Jd[2]):
                s2 = np.zeros(M, ck.dtype)
                for j2 in range(Jd[1]):
                    s1 = np.zeros(M, ck.dtype)
                    base = k3[:, j3] * strides[2] + k2[:, j2] * strides[1]
                    for j1 in range(Jd[0]):
                        s1 += c1[:, j1] * g[base + k1[:, j1]]
                    s2 += c2[:, j2] * s1
                s3 += c3[:, j3] * s2
            out[:, r] = s3
    return out


def interp_adjoint(fm, Kd, hs, Jd, L, tm, cplx_kernel=False):
    """Adjoint table1 scatter, ``nufft_table.template.c:144-305, 308-531, 924-1202``.

    ``ck[kk] += conj(coef) * fm[m]`` (the complex kernel is conjugated,
    ``:189-190, :475-476, :490-491``).  Returns (prod(Kd), nreps); the output
    starts from zero as in ``:336-337``.
    """
    ndim = len(Kd)
    fm = np.asarray(fm)
    if fm.ndim == 1:
        fm = fm[:, None]
    taps = [_axis_taps(tm[:, d], Jd[d], L, Kd[d], hs[d], cplx_kernel)
            for d in range(ndim)]
    strides = [int(np.prod(Kd[:d])) for d in range(ndim)]
    nk = int(np.prod(Kd))
    out = np.zeros((nk, fm.shape[1]), dtype=fm.dtype)
    real_t = np.float32 if fm.dtype == np.complex64 else np.float64
    for r in range(fm.shape[1]):
        y = fm[:, r]
        acc_re = np.zeros(nk, np.float64)
        acc_im = np.zeros(nk, np.float64)
        # Every tap value conj(c) * y is formed in the operator's precision as
        # in the C kernels; the sums accumulate in float64 (bincount), which
        # is within the tolerance of the serial float32 C accumulation.  The
        # taps of one outer-axis index are collected and added by one
        # bincount over the grid.
        idx_l, val_l = [], []

        def add(idx, v):
            idx_l.append(idx)
            val_l.append(v)

        def flush():
            if not idx_l:
                return
            idx = np.concatenate(idx_l)
            v = np.concatenate(val_l)
            acc_re[:] += np.bincount(idx, weights=v.real.astype(np.float64), minlength=nk)
            acc_im[:] += np.bincount(idx, weights=v.imag.astype(np.float64), minlength=nk)
            idx_l.clear()
            val_l.clear()

        if ndim == 1:
            (k1, c1), = taps
            for j1 in range(Jd[0]):
                add(k1[:, j1], np.conj(c1[:, j1]) * y)
            flush()
        elif ndim == 2:
            (k1, c1), (k2, c2) = taps
            for j2 in range(Jd[1]):
                v2 = np.conj(c2[:, j2]) * y
                base = k2[:, j2] * strides[1]
                for j1 in range(Jd[0]):
                    add(base + k1[:, j1], np.conj(c1[:, j1]) * v2)
            flush()
        else:
            (k1, c1), (k2, c2), (k3, c3) = taps
            for j3 in range(Jd[2]):
                v3 = np.conj(c3[:, j3]) * y
                for j2 in range(Jd[1]):
                    v2 = np.conj(c2[:, j2]) * v3
                    base = k3[:, j3] * strides[2] + k2[:, j2] * strides[1]
                    for j1 in range(Jd[0]):
                        add(base + k1[:, j1], np.conj(c1[:, j1]) * v2)
                flush()
        out[:, r] = (acc_re.astype(real_t) + 1j * acc_im.astype(real_t))
    return out


# --------------------------------------------------------------------------
# exact NUDFT (float64)
# --------------------------------------------------------------------------
def _grid_coords(shape, n_shift):
    nn = [np.arange(shape[d]) - n_shift[d] for d in range(len(shape))]
    nn = np.meshgrid(*nn, indexing="ij")
    return [g.ravel(order="F").astype(np.float64) for g in nn]


def dtft(x, omega, shape, n_shift=None, chunk=4096):
    """Exact ``X[m] = sum_n x[n] exp(-i omega_m . (n - n_shift))``.

    ``_dtft.py:16-123`` (dense form ``:115-118``), chunked over m.
    ``x`` is (prod(shape),) or (prod(shape), L) in Fortran order.
    """
    omega = np.asarray(omega, dtype=np.float64)
    dd = omega.shape[1]
    n_shift = np.zeros(dd) if n_shift is None else np.atleast_1d(np.squeeze(n_shift))
    x = np.asarray(x).reshape((int(np.prod(shape)), -1), order="F").astype(np.complex128)
    nn = _grid_coords(shape, n_shift)
    out = np.empty((omega.shape[0], x.shape[1]), np.complex128)
    for s in range(0, omega.shape[0], chunk):
        om = omega[s:s + chunk]
        arg = np.outer(om[:, 0], nn[0])
        for d in range(1, dd):
            arg += np.outer(om[:, d], nn[d])
        out[s:s + chunk] = np.exp(-1j * arg) @ x
    return out[:, 0] if out.shape[1] == 1 else out


def dtft_adj(X, omega, shape, n_shift=None, chunk=4096, pixels=None):
    """Exact adjoint ``x[n] = sum_m X[m] exp(+i omega_m . (n - n_shift))``.

    ``_dtft.py:126-216`` (dense form ``:207-210``), chunked over m.
    Returns (prod(shape),) or (prod(shape), L) in Fortran order; with
    ``pixels`` (Fortran flat indices) only those entries, in that order.
    """
    omega = np.asarray(omega, dtype=np.float64)
    dd = omega.shape[1]
    n_shift = np.zeros(dd) if n_shift is None else np.atleast_1d(np.squeeze(n_shift))
    X = np.asarray(X).reshape((omega.shape[0], -1)).astype(np.complex128)
    nn = _grid_coords(shape, n_shift)
    if pixels is not None:
        nn = [g[np.asarray(pixels)] for g in nn]
    out = np.zeros((len(nn[0]), X.shape[1]), np.complex128)
    for s in range(0, omega.shape[0], chunk):
        om = omega[s:s + chunk]
        arg = np.outer(nn[0], om[:, 0])
        for d in range(1, dd):
            arg += np.outer(nn[d], om[:, d])
        out += np.exp(1j * arg) @ X[s:s + chunk]
    return out[:, 0] if out.shape[1] == 1 else out


# --------------------------------------------------------------------------
# operator composition
# --------------------------------------------------------------------------
class NufftOracle:
    """Restatement of ``NufftBase`` in table mode (``_nufft.py:216-359``).

    ``fft``/``adj`` restate ``nufft_forward`` (``:1275-1397``) and
    ``nufft_adj`` (``:1459-1578``) with Fortran-ordered arrays; the FFT is
    ``numpy.fft`` (the reference's ``mrrt.utils.fftn`` is not vendored, so the
    FFT rounding is not pinned).  ``interp`` selects the table1 kernels:
    ``"numpy"`` (this module) or a callable pair from ``oracle.ref_ctypes``.
    """

    def __init__(self, Nd, omega, Jd=4, Kd=None, precision="single", Ld=1024,
                 ortho=False, n_shift=None, phasing="real",
                 adjoint_scalefactor=1.0, interp=None, mode="table"):
        Nd = (Nd,) if np.isscalar(Nd) else tuple(int(n) for n in Nd)
        ndim = len(Nd)
        Jd = (int(Jd),) * ndim if np.isscalar(Jd) else tuple(int(j) for j in Jd)
        if Kd is None:
            Kd = tuple(int(1.5 * n) for n in Nd)
        Kd = (int(Kd),) * ndim if np.isscalar(Kd) else tuple(int(k) for k in Kd)
        self.Nd, self.Jd, self.Kd, self.Ld, self.ndim = Nd, Jd, Kd, int(Ld), ndim
        self.phasing = phasing
        self.ortho = ortho
        self.scale_ortho = sqrt(float(np.prod(Kd))) if ortho else 1
        self.adjoint_scalefactor = adjoint_scalefactor
        omega = np.asarray(omega)
        if omega.ndim == 1:
            omega = omega[:, None]
        if precision == "auto":
            precision = "single" if omega.dtype == np.float32 else "double"
        self.precision = precision
        self.rdt = np.dtype(np.float32 if precision == "single" else np.float64)
        self.cdt = np.dtype(np.complex64 if precision == "single" else np.complex128)
        self.n_shift = (0.0,) * ndim if n_shift is None else tuple(
            float(s) for s in np.broadcast_to(np.asarray(n_shift, dtype=float).ravel(), (ndim,)))
        if phasing == "real":
            self.n_mid = tuple(n // 2 for n in Nd)
        else:
            self.n_mid = tuple((n - 1) / 2.0 for n in Nd)
        self.alpha = [beatty_alpha(Jd[d], Kd[d], Nd[d]) for d in range(ndim)]
        sn = np.array([1.0])
        for d in range(ndim):
            sn = np.outer(sn.ravel(), calc_scaling_1d(Nd[d], Kd[d], Jd[d], self.alpha[d], self.n_mid[d]))
        self.sn = sn.reshape(Nd).astype(self.rdt)
        if phasing == "real":
            self.phase_before = phase_before(Kd, self.n_mid, self.rdt)
            self.phase_after = phase_after(omega, self.n_mid, self.n_shift, self.rdt)
        else:
            self.phase_before = None
            self.phase_after = None
        self.omega = omega.astype(self.rdt)
        self.M = self.omega.shape[0]
        self.mode = mode
        if mode == "sparse":
            # _init_sparsemat re-runs _set_phase_funcs on the cast omega
            # (:768-769); the complex kernel's shift phase goes into P (:833-838)
            if phasing == "real":
                self.phase_after = phase_after(self.omega, self.n_mid, self.n_shift, self.rdt)
            self.phase_shift = None
            self.p = sparse_matrix(self.omega, Nd, Jd, Kd, self.alpha, phasing, self.n_shift, self.rdt)
            self._fwd = lambda ck: np.asarray(self.p @ ck)          # :1384
            self._adj = lambda fm: np.asarray(self.p.getH() @ fm)   # :1505
            self.h = None
            self.tm = compute_tm(self.omega, Kd)
            return
        if phasing == "complex" and any(s != 0 for s in self.n_shift):
            # _nufft.py:898-901: from the cast omega, kept complex128
            self.phase_shift = np.exp(1j * np.dot(self.omega, np.asarray(self.n_shift)))
        else:
            self.phase_shift = None
        self.h = []
        for d in range(ndim):
            h, _ = table_make1("fast", Nd[d], Jd[d], Kd[d], self.Ld, phasing)
            if phasing == "complex":
                h = h.astype(self.cdt)
            else:
                h = np.real(h).astype(self.rdt)
            self.h.append(h)
        self.tm = compute_tm(self.omega, Kd)
        if interp is None or interp == "numpy":
            self._fwd = lambda ck: interp_forward(ck, self.Kd, self.h, self.Jd, self.Ld, self.tm, phasing == "complex")
            self._adj = lambda fm: interp_adjoint(fm, self.Kd, self.h, self.Jd, self.Ld, self.tm, phasing == "complex")
        else:
            fwd, adj = interp
            self._fwd = lambda ck: fwd(ck, self.Kd, self.h, self.Jd, self.Ld, self.tm)
            self._adj = lambda fm: adj(fm, self.Kd, self.h, self.Jd, self.Ld, self.tm)

    def fft(self, x, grid_only=False):
        """``nufft_forward`` (``_nufft.py:1275-1397``), Fortran order.

        ``grid_only``: ``x`` is the oversampled grid (prod(Kd), reps); only the
        table interpolation (and the complex kernel's phase_shift) is applied
        and the (M, reps) result is returned unsqueezed (:1302-1304, :1376-1388).
        """
        Nd, Kd = self.Nd, self.Kd
        if grid_only:
            xk = np.asfortranarray(x).reshape((int(np.prod(Kd)), -1), order="F").astype(self.cdt)
            y = self._fwd(xk)
            if self.phase_shift is not None:
                y = (y * self.phase_shift[:, None]).astype(self.cdt)
            return y.reshape((self.M, xk.shape[1]), order="F")
        x = np.asfortranarray(x).reshape(list(Nd) + [-1], order="F").astype(self.cdt)
        n_reps = x.shape[-1]
        xk = x * self.sn[..., None]
        xk = np.fft.fftn(xk, Kd, axes=tuple(range(self.ndim))).astype(self.cdt)
        if self.phase_before is not None:
            xk *= self.phase_before[..., None]
        xk = xk.reshape((int(np.prod(Kd)), n_reps), order="F")
        if self.ortho:
            xk /= self.scale_ortho
        y = self._fwd(xk)
        if self.phase_shift is not None:          # _nufft.py:1086-1095
            y = (y * self.phase_shift[:, None]).astype(self.cdt)
        y = y.reshape((self.M, n_reps), order="F")
        if self.phase_after is not None:
            y *= self.phase_after[:, None]
        return y[:, 0] if n_reps == 1 else y

    def adj(self, y, return_psf=False, grid_only=False):
        """``nufft_adj`` (``_nufft.py:1459-1578``), Fortran order.

        ``grid_only``: return the raw gridded samples (prod(Kd), reps) before
        conj(phase_before) and the FFT (:1507-1509); ``return_psf``: skip
        conj(phase_after) and return the first repetition's grid as Kd
        (:1495, :1517-1518).
        """
        Nd, Kd = self.Nd, self.Kd
        y = np.asfortranarray(y).astype(self.cdt).reshape((self.M, -1), order="F").copy()
        n_reps = y.shape[-1]
        if self.phase_after is not None and not return_psf:
            y *= self.phase_after.conj()[:, None]
        if self.phase_shift is not None:          # _nufft.py:1141-1150
            y = (y * self.phase_shift.conj()[:, None]).astype(self.cdt)
        g = self._adj(y)
        if grid_only:
            return g
        g = g.reshape(Kd + (n_reps,), order="F")
        if return_psf:
            return g[..., 0]
        if self.phase_before is not None:
            g *= self.phase_before.conj()[..., None]
        x = np.fft.ifftn(g, Kd, axes=tuple(range(self.ndim))).astype(self.cdt)
        x = x[tuple([slice(n) for n in Nd] + [slice(None)])]
        if self.ortho:
            x *= self.scale_ortho * self.adjoint_scalefactor
        else:
            x *= np.prod(Kd) * self.adjoint_scalefactor
        x *= np.conj(self.sn)[..., None]
        return x[..., 0] if n_reps == 1 else x
