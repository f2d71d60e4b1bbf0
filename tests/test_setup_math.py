"""Product set-up math (mrrt/nufft_amd/_kernel_math.py) against the oracle."""

import numpy as np
import pytest

from mrrt.nufft_amd import _kernel_math as P
from oracle import nufft_oracle as O


@pytest.mark.parametrize("N", [16, 17, 64, 65, 128, 256])
@pytest.mark.parametrize("J", [4, 5, 6, 7, 8])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_build_table_bit_exact(N, J, phasing):
    """Direct table builder == the oracle's sparse-matrix route, bit for bit."""
    for K in (2 * N, 2 * N + 1, int(1.5 * N)):
        for L in (256, 1024):
            a = O.table_make1("fast", N, J, K, L, phasing)[0]
            b = P.build_table(N, J, K, L, phasing)
            assert a.dtype == b.dtype
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), (N, J, K, L, phasing)


@pytest.mark.parametrize("N,K,J", [(128, 256, 6), (256, 512, 6), (16, 33, 7), (65, 130, 6), (8, 16, 6)])
def test_deapod_factors(N, K, J):
    al = P.beatty_alpha(J, K, N)
    assert al == O.beatty_alpha(J, K, N)
    for n_mid in (N // 2, (N - 1) / 2.0):
        a = O.calc_scaling_1d(N, K, J, al, n_mid)
        b = P.deapod_factors(N, K, J, al, n_mid)
        np.testing.assert_allclose(b, a, rtol=1e-13)


def test_kb_ft_branches():
    """Both branches of the closed form agree with the oracle's Bessel route."""
    J, al = 6, 5.0
    u = np.linspace(-0.6, 0.6, 121)
    np.testing.assert_allclose(P.kaiser_bessel_ft(u, J, al), O.kaiser_bessel_ft(u, J, al), rtol=1e-11, atol=1e-14)


def test_kaiser_bessel_same_as_oracle():
    J, al = 6, P.beatty_alpha(6, 512, 256)
    x = np.linspace(-3.5, 3.5, 1001)
    np.testing.assert_array_equal(P.kaiser_bessel(x, J, al), O.kaiser_bessel(x, J, al))
