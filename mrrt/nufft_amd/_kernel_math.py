"""Set-up math of the table1 NUFFT: KB kernel, its Fourier transform, the
Beatty shape parameter, the interpolation tables and the deapodization
factors.  Host-side (numpy/scipy), evaluated once per plan, as in the
reference.

The interpolation table reproduces the reference's values bit for bit: the
reference builds it through a float32 1-D sparse NUFFT
(mrrt/nufft/_nufft.py:1195-1243 with the default precision 'single' of
:222), i.e. float32-rounded sample positions, a float64 KB evaluation
(_utils.py:47-86, _kaiser_bessel.py:78-149) and float32 storage
(_nufft.py:843-858).  ``build_table`` evaluates the same float32/float64
expressions directly, without forming the sparse matrix.
"""

import warnings
from math import sqrt

import numpy as np
from scipy.special import gamma, i0, iv, jv

__all__ = [
    "beatty_alpha",
    "kaiser_bessel",
    "kaiser_bessel_ft",
    "build_table",
    "deapod_factors",
]


def beatty_alpha(J, K, N):
    """KB shape parameter of Beatty et al. 2005, eq. 5 (_kernels.py:151-154; m = 0 at :160)."""
    k_n = K / N
    return np.pi * sqrt(J ** 2 / k_n ** 2 * (k_n - 0.5) ** 2 - 0.8)


def kaiser_bessel(x=None, J=6, alpha=None, m=0, K_N=None):
    """Generalized Kaiser-Bessel window on |x| < J/2, zero elsewhere
    (_kaiser_bessel.py:78-149; Lewitt, JOSA-A 1990, eq. A1):

        f^m I_m(alpha f) / I_m(alpha),  f = sqrt(1 - (2x/J)^2).

    ``alpha`` defaults to 2.34 J.  A negative non-integer order uses |m| in
    the Bessel functions, with a warning, as the reference does.  ``K_N``
    (grid oversampling) is accepted for the reference's signature and not
    used.  The order-0 window is the one every NUFFT table uses.
    """
    if alpha is None:
        alpha = 2.34 * J
    if m < 0 and abs(round(m) - m) > np.finfo(float).eps:
        warnings.warn("Negative NonInt m=%g\n\t- using modified definition of KB function\n" % m)
    m_bi = abs(m)
    x = np.asarray(x)
    inside = 2 * np.abs(x) < J
    u = 2 * x[inside] / J
    u *= u
    f = np.sqrt(1 - u)
    out = np.zeros_like(x)
    if m_bi == 0:
        out[inside] = i0(alpha * f) / float(i0(alpha))
    else:
        out[inside] = (f ** m * iv(m_bi, alpha * f)) / float(iv(m_bi, alpha))
    return out.real


def kaiser_bessel_ft(u, J=6, alpha=None, m=0, d=1):
    """Fourier transform of the generalized KB window in d dimensions
    (_kaiser_bessel.py:153-226; Lewitt, JOSA-A 1990, eq. A3).

    Order 0, d = 1 (every NUFFT's deapodization): the closed form
    J sinh(w) / (I0(alpha) w), w = sqrt(alpha^2 - (pi J u)^2);
    J sin(z) / (I0(alpha) z), z = sqrt((pi J u)^2 - alpha^2) past the
    transition; J / I0(alpha) at w = 0.  Otherwise
    (2 pi)^(d/2) (J/2)^d alpha^m / I_m(alpha) J_nu(z) / z^nu, nu = d/2 + m,
    z = sqrt((pi J u)^2 - alpha^2) taken complex where the radicand is
    negative.
    """
    if not alpha:
        alpha = 2.34 * J
    u = np.asarray(u, dtype=np.float64)
    if m == 0 and d == 1:
        q = (np.pi * J) * u
        q = alpha * alpha - q * q
        out = np.empty_like(q)
        pos = q > 0
        neg = q < 0
        w = np.sqrt(q[pos])
        out[pos] = np.sinh(w) / w
        z = np.sqrt(-q[neg])
        out[neg] = np.sin(z) / z
        out[~(pos | neg)] = 1.0
        return J * out / i0(alpha)
    if m < -1:
        warnings.warn("m=%g < -1 in kaiser_bessel_ft(): validity of the FT formula uncertain" % m)
    elif m < 0 and abs(round(m) - m) > np.finfo(float).eps:
        warnings.warn("negative non-integer m=%g in kaiser_bessel_ft(): validity of the FT formula uncertain" % m)
    q = (np.pi * J) * u
    z = np.lib.scimath.sqrt(q * q - alpha * alpha)
    nu = d / 2.0 + m
    const = (2 * np.pi) ** (d / 2.0) * (J / 2.0) ** d * alpha ** m / (iv(m, alpha) if m else i0(alpha))
    with np.errstate(invalid="ignore", divide="ignore"):
        y = const * jv(nu, z) / z ** nu
    # z = 0 (the transition point): the limit of J_nu(z) / z^nu is 2^-nu / Gamma(nu + 1)
    zero = z == 0
    if np.any(zero):
        y = np.where(zero, const * 2.0 ** -nu / gamma(nu + 1), y)
    return np.real(y)


def build_table(N, J, K, L, phasing="real"):
    """Table1 interpolation table h[0..J*L] for one axis.

    Same values as the reference's _nufft_table_make1(how='fast')
    (_nufft.py:1223-1243): row m of the 1-D sparse matrix holds the J
    coefficients of the sample t1[m] = J/2 - 1 + m/L (m + 1 for odd N),
    taken at float32 positions; column c (< J) of that row lands at
    h[(J-1-c)*L + m]; the wrap entry repeats the first (even N) or last (odd N)
    value.  Returns float32 (real phasing) or complex64 (complex phasing).
    """
    alpha = beatty_alpha(J, K, N)
    if N % 2 == 0:
        t1 = J / 2.0 - 1 + np.arange(L) / L
    else:
        t1 = J / 2.0 - 1 + np.arange(1, L + 1) / L
    om = (t1 * 2 * np.pi / K).astype(np.float32)      # omega of the nested operator, cast at :315
    gam = 2 * np.pi / K
    u = om / gam                                       # float32 (_utils.py:41, :75)
    koff = np.floor(u - J / 2.0)                       # float32
    dk = u - koff                                      # float32
    jj = np.arange(1, J + 1)
    arg = -jj[:, None] + dk[None, :]                   # float64 (J, L)
    coef = kaiser_bessel(arg, J, alpha)
    if phasing == "complex":
        phase_scale = 1j * gam * (N - 1) / 2.0
        coef = (np.exp(phase_scale * arg) * coef).conj()
        vals = coef.astype(np.complex64)
    else:
        vals = coef.astype(np.float32)
    cols = np.mod(jj[:, None] + koff[None, :].astype(np.int64), K)   # (J, L)
    table = np.zeros((J, L), dtype=vals.dtype)        # table[J-1-c, m] = p[m, c]
    rows = np.broadcast_to(np.arange(L)[None, :], cols.shape)
    keep = cols < J
    np.add.at(table, (J - 1 - cols[keep], rows[keep]), vals[keep])
    h = table.ravel()
    if N % 2 == 0:
        h = np.concatenate((h, h[:1]))
    else:
        h = np.concatenate((h[-1:], h))
    return h


def deapod_factors(N, K, J, alpha, n_mid):
    """1 / KB_FT((n - n_mid) / K), n = 0..N-1, float64 (_nufft.py:737-747)."""
    nc = np.arange(-n_mid, -n_mid + N)
    return 1.0 / kaiser_bessel_ft(nc / K, J, alpha)
