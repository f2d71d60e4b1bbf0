"""Public names and signatures of the drop-in (mrrt/nufft/__init__.py:1-15):
kernel classes and the generalized Kaiser-Bessel functions.  No GPU needed.

The order-0 window and transform are checked against the oracle; the
order-m transforms against numerical quadrature of the window itself
(Lewitt 1990, eq. A3, restated in _kernel_math.kaiser_bessel_ft)."""

import inspect

import numpy as np
import pytest
from scipy.integrate import quad

import mrrt.nufft_amd as M
from mrrt.nufft_amd import _kernel_math as K
from oracle import nufft_oracle as O


def test_public_names():
    for name in ("dtft", "dtft_adj", "BeattyKernel", "KaiserBesselKernel", "NufftBase", "kaiser_bessel",
                 "kaiser_bessel_ft", "nufft_forward", "nufft_adj"):
        assert hasattr(M, name), name
        assert name in M.__all__


def test_reference_signatures():
    """Parameter names and order of the reference's functions."""
    sig = lambda f: list(inspect.signature(f).parameters)  # noqa: E731
    assert sig(M.nufft_forward)[:5] == ["obj", "x", "copy_x", "grid_only", "xp"]
    assert sig(M.nufft_adj)[:6] == ["obj", "xk", "copy", "return_psf", "grid_only", "xp"]
    assert sig(M.dtft)[:6] == ["x", "omega", "shape", "n_shift", "useloop", "xp"]
    assert sig(M.dtft_adj)[:6] == ["xk", "omega", "shape", "n_shift", "useloop", "xp"]
    assert sig(M.kaiser_bessel)[:5] == ["x", "J", "alpha", "m", "K_N"]
    assert sig(M.kaiser_bessel_ft)[:5] == ["u", "J", "alpha", "m", "d"]


def test_beatty_kernel():
    k = M.BeattyKernel((6, 4), (128, 100), (256, 150))
    assert k.shape == (6, 4) and k.ndim == 2 and k.kernel_type == "kb:beatty" and k.m == [0, 0]
    assert k.alpha[0] == pytest.approx(O.beatty_alpha(6, 256, 128), rel=1e-15)
    assert k.alpha[1] == pytest.approx(O.beatty_alpha(4, 150, 100), rel=1e-15)
    x = np.linspace(-3.5, 3.5, 71)
    np.testing.assert_array_equal(k.kernels[0](x), O.kaiser_bessel(x, 6, k.alpha[0]))
    assert "kb:beatty" in str(k)


def test_kaiser_bessel_kernel_integer_params():
    """As in the reference (_as_1d_ints, _utils.py:89-127): alpha and m must
    be integer-valued; scalars broadcast to every axis."""
    k = M.KaiserBesselKernel((6, 6), alpha=14, m=0)
    assert k.alpha == [14, 14] and k.m == [0, 0] and k.kernel_type == "kb:user"
    x = np.linspace(-3, 3, 31)
    np.testing.assert_allclose(k.kernels[1](x), O.kaiser_bessel(x, 6, 14.0), rtol=1e-15, atol=0)
    with pytest.raises(ValueError):
        M.KaiserBesselKernel(6, alpha=13.5, m=0)
    with pytest.raises(ValueError):
        M.KaiserBesselKernel(6, alpha=None, m=0)
    with pytest.raises(ValueError):
        M.KaiserBesselKernel((6, 6), alpha=[1, 2, 3], m=0)


def test_order_zero_unchanged():
    a = K.beatty_alpha(6, 512, 256)
    x = np.linspace(-3.2, 3.2, 101)
    np.testing.assert_array_equal(M.kaiser_bessel(x, 6, a), O.kaiser_bessel(x, 6, a))
    u = np.linspace(-0.3, 0.3, 13)
    np.testing.assert_allclose(M.kaiser_bessel_ft(u, 6, a), O.kaiser_bessel_ft(u, 6, a), rtol=1e-12)


@pytest.mark.parametrize("m", [1, 2])
def test_generalized_kb_ft_vs_quadrature(m):
    """1-D transform of the order-m window against quadrature of the window."""
    J, a = 6, K.beatty_alpha(6, 512, 256)
    u = np.linspace(-0.3, 0.3, 7)
    ft = M.kaiser_bessel_ft(u, J, a, m=m)
    q = np.array([quad(lambda t: M.kaiser_bessel(np.array([t]), J, a, m=m)[0] * np.cos(2 * np.pi * v * t),
                       -J / 2, J / 2, limit=200)[0] for v in u])
    np.testing.assert_allclose(ft, q, rtol=1e-9, atol=1e-12 * np.abs(q).max())


def test_kb_ft_3d_radial_vs_quadrature():
    """d = 3: the transform of the radial window is 4 pi / (2 pi u) int r f(r) sin(2 pi u r) dr."""
    J, a = 4, 10.0
    for u in (0.05, 0.2):
        q = quad(lambda r: r * M.kaiser_bessel(np.array([r]), J, a)[0] * np.sin(2 * np.pi * u * r), 0, J / 2)[0]
        ref = 4 * np.pi / (2 * np.pi * u) * q
        assert M.kaiser_bessel_ft(np.array([u]), J, a, m=0, d=3)[0] == pytest.approx(ref, rel=1e-9)
