"""Pins the CPU oracle (oracle/nufft_oracle.py) before it is trusted.

Each check compares the oracle with something other than its own formula:
the reference's known-answer tests and test cases, the reference's own C
kernels (golden vectors written by tests/golden/make_golden.py), closed
forms, numerical quadrature, and brute force on tiny inputs.
"""

import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import i0

from oracle import nufft_oracle as O
from oracle import ref_ctypes as R
from tests.cases import golden_cases, perturbed_gridpoints, rel_l2, uniform_freqs


# ---------------------------------------------------------------- set-up math
def test_nufft_offset_kat():
    """mrrt/nufft/tests/test_utils.py:18-22."""
    assert O.nufft_offset(0, 4, 128) == -2
    assert O.nufft_offset(0, 5, 128) == -3
    assert O.nufft_offset(0, 5.5, 128) == -3
    assert O.nufft_offset(0, 6, 128) == -3
    np.testing.assert_equal(O.nufft_offset(np.asarray([0]), 6.01, 128), np.asarray([-4]))


def test_beatty_alpha_value():
    # alpha(J=6, K/N=2) = pi*sqrt(36/4*1.5^2 - 0.8) (Beatty 2005 eq. 5)
    assert abs(O.beatty_alpha(6, 512, 256) - 13.855100346) < 1e-8
    assert abs(O.beatty_alpha(6, 256, 128) - np.pi * np.sqrt(9 * 2.25 - 0.8)) < 1e-12


def test_kaiser_bessel_support_and_peak():
    """Support properties of mrrt/nufft/tests/test_kernels.py:56-70."""
    for J in (3, 4, 6):
        a = O.beatty_alpha(J, 128, 64)
        x = np.linspace(-J / 2 + 0.001, J / 2 - 0.001, 100)
        assert np.all(O.kaiser_bessel(x, J, a) > 0)
        for v in (J / 2, J / 2 + 10, -J / 2, -J / 2 - 10):
            assert O.kaiser_bessel(np.asarray([v]), J, a)[0] == 0
        assert O.kaiser_bessel(np.asarray([0.0]), J, a)[0] == 1.0
        x = np.linspace(-J / 2, J / 2, 257)
        np.testing.assert_array_equal(O.kaiser_bessel(x, J, a), O.kaiser_bessel(-x, J, a))


def test_kaiser_bessel_series():
    """I0 by its power series sum (z/2)^(2k)/(k!)^2, independent of scipy."""
    J, a = 6, O.beatty_alpha(6, 512, 256)
    x = np.linspace(-2.9, 2.9, 41)

    def i0_series(z):
        s, term, k = 0.0, 1.0, 0
        while term > 1e-18 * max(s, 1):
            s += term
            k += 1
            term *= (z / 2) ** 2 / k ** 2
        return s

    ref = np.array([i0_series(a * np.sqrt(1 - (2 * v / J) ** 2)) / i0_series(a) for v in x])
    np.testing.assert_allclose(O.kaiser_bessel(x, J, a), ref, rtol=1e-13)


@pytest.mark.parametrize("J,KN", [(6, 2.0), (4, 2.0), (7, 1.5)])
def test_kaiser_bessel_ft_quadrature(J, KN):
    """KB_FT(u) = int KB(x) exp(-i 2 pi u x) dx over [-J/2, J/2]."""
    a = O.beatty_alpha(J, int(64 * KN), 64)
    for u in (0.0, 0.05, 0.13, 0.25, 0.4):
        val, _ = quad(lambda x: O.kaiser_bessel(np.asarray([x]), J, a)[0] * np.cos(2 * np.pi * u * x),
                      -J / 2, J / 2, limit=200, epsabs=1e-14, epsrel=1e-13)
        assert abs(O.kaiser_bessel_ft(np.asarray([u]), J, a)[0] - val) < 1e-10 * abs(val)


@pytest.mark.parametrize("n,phasing", [(64, "real"), (65, "real"), (64, "complex"), (65, "complex")])
def test_table_make1_slow_fast_ratio(n, phasing):
    """mrrt/nufft/tests/test_nufft.py:391-409."""
    h0, t0 = O.table_make1("slow", n, 6, 2 * n, 2048, phasing)
    h1, t1 = O.table_make1("fast", n, 6, 2 * n, 2048, phasing)
    h2, t2 = O.table_make1("ratio", n, 6, 2 * n, 2048, phasing)
    np.testing.assert_array_almost_equal(h0, h1, decimal=6)
    np.testing.assert_array_almost_equal(h0, h2, decimal=6)
    np.testing.assert_array_almost_equal(t0, t1, decimal=6)
    np.testing.assert_array_almost_equal(t0, t2, decimal=6)


@pytest.mark.parametrize("N,J,K,L", [(256, 6, 512, 1024), (128, 6, 256, 512), (16, 7, 33, 1024), (65, 6, 130, 256)])
def test_table_vs_closed_form(N, J, K, L):
    """h[i] = KB(-J/2 + i/L) up to the float32 rounding of position and value."""
    h, t0 = O.table_make1("fast", N, J, K, L, "real")
    assert h.dtype == np.float32 and h.size == J * L + 1
    a = O.beatty_alpha(J, K, N)
    exact = O.kaiser_bessel(t0, J, a)
    # slope of KB <= ~alpha; position error <= ulp32(J) ~ 5e-7
    assert np.max(np.abs(h - exact)) < 3e-6
    if N % 2 == 0:
        assert h[-1] == h[0]


def test_table_complex_phase():
    """Complex table = KB(t) exp(-i pi (N-1) t / K) (conjugated ud, _nufft.py:793-796, :830-831)."""
    N, J, K, L = 64, 6, 128, 512
    h, t0 = O.table_make1("fast", N, J, K, L, "complex")
    a = O.beatty_alpha(J, K, N)
    exact = O.kaiser_bessel(t0, J, a) * np.exp(-1j * np.pi * (N - 1) * t0 / K)
    assert np.max(np.abs(h - exact)) < 3e-6


# ------------------------------------------------------------------ hot loops
@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["key"])
def test_interp_vs_reference_golden(case):
    """Oracle gather/scatter == the reference's C kernels on the golden vectors."""
    kw = dict(Kd=case["Kd"], hs=case["hs"], Jd=case["Jd"], L=case["L"], tm=case["tm"],
              cplx_kernel=case["complex_kernel"])
    fwd = O.interp_forward(case["ck"], **kw)
    adj = O.interp_adjoint(case["fm"], **kw)
    single = case["tm"].dtype == np.float32
    tol = 2e-6 if single else 1e-14
    assert rel_l2(fwd, case["fwd"]) <= tol
    assert rel_l2(adj, case["adj"]) <= tol
    if not case["complex_kernel"]:
        # real-kernel forward: same operations in the same order as the C code
        np.testing.assert_array_equal(fwd, case["fwd"])


@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_interp_adjointness(ndim):
    rng = np.random.default_rng(ndim)
    Kd = (20, 18, 16)[:ndim]
    Nd = tuple(k // 2 for k in Kd)
    Jd = (6, 5, 4)[:ndim]
    L = 512
    for ph in ("real", "complex"):
        hs = [O.table_make1("fast", Nd[d], Jd[d], Kd[d], L, ph)[0].astype(np.complex128 if ph == "complex" else np.float64)
              for d in range(ndim)]
        tm = rng.uniform(-10, 10, (200, ndim))
        x = rng.standard_normal((int(np.prod(Kd)), 1)) + 1j * rng.standard_normal((int(np.prod(Kd)), 1))
        y = rng.standard_normal((200, 1)) + 1j * rng.standard_normal((200, 1))
        Ax = O.interp_forward(x, Kd, hs, Jd, L, tm, ph == "complex")
        Ahy = O.interp_adjoint(y, Kd, hs, Jd, L, tm, ph == "complex")
        lhs = np.vdot(y, Ax)
        rhs = np.vdot(Ahy, x)
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_interp_bruteforce_tiny():
    """Table1 gather == sum over the J^2 nearest grid points of the float64
    closed-form KB, up to the linear-interpolation error of the table."""
    rng = np.random.default_rng(5)
    Kd, Nd, Jd, L = (16, 14), (8, 7), (6, 6), 1024
    hs = [O.table_make1("fast", Nd[d], Jd[d], Kd[d], L, "real")[0].astype(np.float64) for d in range(2)]
    al = [O.beatty_alpha(Jd[d], Kd[d], Nd[d]) for d in range(2)]
    tm = rng.uniform(-20, 20, (30, 2))
    g = rng.standard_normal(int(np.prod(Kd))) + 1j * rng.standard_normal(int(np.prod(Kd)))
    got = O.interp_forward(g[:, None], Kd, hs, Jd, L, tm)[:, 0]
    G = g.reshape(Kd, order="F")
    for m in range(tm.shape[0]):
        acc = 0
        for k1 in range(int(np.floor(tm[m, 0])) - 4, int(np.floor(tm[m, 0])) + 5):
            for k2 in range(int(np.floor(tm[m, 1])) - 4, int(np.floor(tm[m, 1])) + 5):
                w = O.kaiser_bessel(np.asarray([tm[m, 0] - k1]), 6, al[0])[0] * \
                    O.kaiser_bessel(np.asarray([tm[m, 1] - k2]), 6, al[1])[0]
                acc += w * G[k1 % Kd[0], k2 % Kd[1]]
        assert abs(got[m] - acc) < 1e-5 * max(1.0, abs(acc))


# ---------------------------------------------------------------- exact NUDFT
@pytest.mark.parametrize("Nd,n_shift", [((16,), [5]), ((8, 12), [1, 3]), ((8, 12, 10), [1, 3, 2])])
def test_dtft_vs_fft(Nd, n_shift):
    """mrrt/nufft/tests/test_dtft.py:39-115."""
    rs = np.random.RandomState(1234)
    x = rs.standard_normal(Nd)
    om = uniform_freqs(Nd)
    Xd = O.dtft(x.ravel(order="F"), om, Nd, n_shift)
    Xf = np.fft.fftn(x).ravel(order="F") * np.exp(1j * (om @ np.asarray(n_shift, float)))
    np.testing.assert_allclose(Xd, Xf, atol=1e-12)


@pytest.mark.parametrize("Nd,n_shift", [((16,), [2]), ((4, 6), [2, 1]), ((32, 16, 2), [2, 1, 3])])
def test_dtft_adj_vs_ifft(Nd, n_shift):
    """mrrt/nufft/tests/test_dtft.py:118-212."""
    rs = np.random.RandomState(1234)
    X = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    om = uniform_freqs(Nd)
    xd = O.dtft_adj(X.ravel(order="F"), om, Nd, n_shift).reshape(Nd, order="F")
    Xp = X * np.exp(-1j * (om @ np.asarray(n_shift, float))).reshape(Nd, order="F")
    xf = np.fft.ifftn(Xp) * np.prod(Nd)
    np.testing.assert_allclose(xd, xf, atol=1e-7)
    # the pixel-subset form picks the same entries (Fortran flat indices)
    pix = np.random.default_rng(3).choice(int(np.prod(Nd)), min(7, int(np.prod(Nd))), replace=False)
    xs = O.dtft_adj(X.ravel(order="F"), om, Nd, n_shift, pixels=pix)
    np.testing.assert_allclose(xs, xf.ravel(order="F")[pix], atol=1e-7)


# ---------------------------------------------------------- composed operator
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_oracle_nufft_adj_kat(phasing):
    """mrrt/nufft/tests/test_nufft.py:60-96 (table mode)."""
    o1 = 2 * np.pi * np.array([0.0, 0.1, 0.3, 0.4, 0.7, 0.9])
    om = np.stack((o1, o1[::-1].copy()), axis=-1)
    st = O.NufftOracle((4, 8), om, Jd=8, Kd=(8, 16), Ld=1024, n_shift=[2.7, 3.1], phasing=phasing)
    data = np.arange(1, 7) ** 2
    xd = O.dtft_adj(data, om, (4, 8), [2.7, 3.1]).reshape((4, 8), order="F")
    xn = st.adj(np.stack((data,) * 3, axis=-1))
    np.testing.assert_array_almost_equal(xd, xn[:, :, -1], decimal=4)


@pytest.mark.parametrize("mode", ["table", "sparse"])
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("Kd,Jd", [((32, 32), 6), ((33, 31), 7)])
def test_oracle_nufft_2d_vs_dtft(precision, phasing, Kd, Jd, mode):
    """mrrt/nufft/tests/test_nufft.py:172-244 (order F)."""
    Nd = (16, 16)
    n_shift = np.asarray(Nd) / 2
    om = perturbed_gridpoints(Nd)
    rs = np.random.RandomState(1234)
    A = O.NufftOracle(Nd, om, Jd=Jd, Kd=Kd, n_shift=n_shift, Ld=1024, precision=precision, phasing=phasing,
                      mode=mode)
    x = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    y = A.fft(x)
    np.testing.assert_allclose(y, O.dtft(x.ravel(order="F"), om, Nd, n_shift), rtol=1e-3, atol=1e-5)
    xa = A.adj(y)
    np.testing.assert_allclose(xa, O.dtft_adj(y, om, Nd, n_shift).reshape(Nd, order="F"), rtol=1e-3, atol=1e-5)


@pytest.mark.parametrize("mode", ["table", "sparse"])
@pytest.mark.parametrize("phasing", ["real", "complex"])
def test_oracle_nufft_1d_3d_vs_dtft(phasing, mode):
    """mrrt/nufft/tests/test_nufft.py:99-169 and :247-324."""
    om = perturbed_gridpoints(64)
    rs = np.random.RandomState(1234)
    A = O.NufftOracle(64, om, Jd=6, Kd=128, n_shift=32, Ld=1024, precision="double", phasing=phasing, mode=mode)
    x = rs.randn(64) + 1j * rs.randn(64)
    y = A.fft(x)
    np.testing.assert_allclose(y, O.dtft(x, om, (64,), [32]), rtol=1e-3, atol=1e-5)
    np.testing.assert_allclose(A.adj(y), O.dtft_adj(y, om, (64,), [32]), rtol=1e-3, atol=1e-5)
    Nd = (8, 8, 8)
    om = perturbed_gridpoints(Nd)
    A = O.NufftOracle(Nd, om, Jd=6, Kd=(16, 16, 16), n_shift=np.asarray(Nd) / 2, Ld=1024,
                      precision="single", phasing=phasing, mode=mode)
    x = rs.standard_normal(Nd) + 1j * rs.standard_normal(Nd)
    y = A.fft(x)
    np.testing.assert_allclose(y, O.dtft(x.ravel(order="F"), om, Nd, np.asarray(Nd) / 2), rtol=1e-2, atol=1e-4)
    np.testing.assert_allclose(A.adj(y), O.dtft_adj(y, om, Nd, np.asarray(Nd) / 2).reshape(Nd, order="F"),
                               rtol=1e-2, atol=1e-4)


@pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (oracle/build_ref.sh)")
@pytest.mark.parametrize("precision,tol", [("single", 1e-6), ("double", 1e-14)])
def test_oracle_composition_with_reference_kernels(precision, tol):
    """The composed oracle gives the same transform with the reference's C
    kernels in place of the numpy hot loops."""
    Nd, Kd = (16, 16), (32, 32)
    om = perturbed_gridpoints(Nd)
    x = np.random.RandomState(3).standard_normal(Nd) + 0.5j
    A = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, precision=precision)
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, precision=precision, interp=(R.ref_interp_forward, R.ref_interp_adjoint))
    y = A.fft(x)
    assert rel_l2(B.fft(x), y) <= tol
    assert rel_l2(B.adj(y), A.adj(y)) <= tol


def test_phase_before_single_quirk():
    """phase_before in float32 carries the float32 rounding of its argument
    (_nufft.py:703-715); in float64 it equals i^(k1+k2) for K = 2N."""
    pb32 = O.phase_before((512, 512), (128, 128), np.float32)
    pb64 = O.phase_before((512, 512), (128, 128), np.float64)
    k = np.arange(512)
    exact = (1j ** ((k[:, None] + k[None, :]) % 4)).astype(np.complex128)
    assert np.max(np.abs(pb64 - exact)) < 1e-10
    err = np.angle(pb32 / exact)
    assert 5e-6 < np.sqrt(np.mean(err ** 2)) < 1e-4   # the quirk is present, and bounded


@pytest.mark.parametrize("N,J,K,L", [(256, 6, 512, 1024), (128, 6, 256, 512), (65, 6, 130, 256)])
def test_table_float32_position_quirk(N, J, K, L):
    """The table is KB at float32-rounded positions (the nested float32
    NufftBase of _nufft.py:1198-1243: omega cast at :315, t = omega / gamma in
    float32 at _utils.py:41, :75), not the float32 rounding of KB at the
    exact positions (which a restatement without the quirk computes exactly:
    t0 = -J/2 + i/L is exact in float64).  The quirk moves about a quarter of
    the entries, each by one or two float32 ulps of 1 (the slope of KB times
    the float32 rounding of the position)."""
    hq, t0 = O.table_make1("fast", N, J, K, L, "real")
    a = O.beatty_alpha(J, K, N)
    naive = O.kaiser_bessel(t0, J, a).astype(np.float32)
    d = np.abs(hq.astype(np.float64) - naive.astype(np.float64))
    ulp = float(np.spacing(np.float32(1.0)))
    assert 0.15 < np.mean(d > 0) < 0.5           # present
    assert ulp / 2 <= d.max() <= 4 * ulp         # and bounded


def test_tm_after_cast_quirk():
    """tm = fl32(fl32(omega) / fl32(2 pi / K)) in single precision (_nufft.py:315
    then :338-342), not fl32(omega / (2 pi / K)) in float64: the two differ in
    the last bits of many samples, by at most 2 ulps of tm."""
    K = 512
    om = np.random.default_rng(4).uniform(-np.pi, np.pi, (100000, 1))
    tq = O.compute_tm(om.astype(np.float32), (K,))[:, 0]
    naive = (om[:, 0] / (2 * np.pi / K)).astype(np.float32)
    assert tq.dtype == np.float32
    diff = np.abs(tq.astype(np.float64) - naive.astype(np.float64))
    assert np.mean(diff > 0) > 0.2
    assert np.all(diff <= 2 * np.spacing(np.abs(naive)))


# ------------------------------------------------------------- sparse mode
@pytest.mark.parametrize("phasing", ["real", "complex"])
@pytest.mark.parametrize("ndim", [1, 2, 3])
def test_sparse_matrix_bruteforce_tiny(phasing, ndim):
    """Every entry of the sparse matrix (_nufft.py:751-877) against a sum
    over the whole grid, without offsets or tap order:
    P[m, k] = prod_d sum_n g_d(t_d - k_d - n K_d), t_d = omega_d / (2 pi / K_d),
    g = KB on |x| < J/2 (times exp(-i pi (N-1) x / K), the conjugated phase of
    :793-796 / :830-831, for the complex kernel), times exp(i omega . n_shift)
    (complex kernel, :833-838).  k is the Fortran flat index."""
    Nd, Kd, Jd, n_shift = [(5, 4, 3), (8, 7, 6), (4, 3, 5), (1.5, -0.5, 2.0)]
    Nd, Kd, Jd, n_shift = Nd[:ndim], Kd[:ndim], Jd[:ndim], n_shift[:ndim]
    M = 6
    om = np.random.default_rng(5).uniform(-np.pi, np.pi, (M, ndim))
    alpha = [O.beatty_alpha(Jd[d], Kd[d], Nd[d]) for d in range(ndim)]
    P = O.sparse_matrix(om, Nd, Jd, Kd, alpha, phasing, n_shift, np.float64).toarray()
    assert P.shape == (M, int(np.prod(Kd)))

    def g(d, x):
        if not 2 * abs(x) < Jd[d]:
            return 0.0
        v = i0(alpha[d] * np.sqrt(1 - (2 * x / Jd[d]) ** 2)) / i0(alpha[d])
        if phasing == "complex":
            v = v * np.exp(-1j * np.pi * (Nd[d] - 1) * x / Kd[d])
        return v

    for m in range(M):
        for k in range(int(np.prod(Kd))):
            kk = np.unravel_index(k, Kd, order="F")
            v = 1.0 + 0j
            for d in range(ndim):
                t = om[m, d] / (2 * np.pi / Kd[d])
                v *= sum(g(d, t - kk[d] - n * Kd[d]) for n in range(-2, 3))
            if phasing == "complex":
                v *= np.exp(1j * np.dot(om[m], np.asarray(n_shift)))
            assert abs(P[m, k] - v) <= 1e-12 * max(1.0, abs(v)), (m, k, P[m, k], v)


def test_sparse_matrix_single_precision_quirks():
    """float32 operator: the matrix is built from the cast omega and stored
    as float32 (:315, :843, :854-858); phase_after then comes from the cast
    omega too (:768-769), unlike table mode (:313)."""
    Nd, Kd = (16, 16), (32, 32)
    om = perturbed_gridpoints(Nd)
    A = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, precision="single", mode="sparse")
    B = O.NufftOracle(Nd, om, Jd=6, Kd=Kd, precision="single", mode="table")
    assert A.p.dtype == np.float32
    alpha = [O.beatty_alpha(6, 32, 16)] * 2
    P64 = O.sparse_matrix(om.astype(np.float32), Nd, (6, 6), Kd, alpha, "real", (0, 0), np.float64)
    np.testing.assert_array_equal(A.p.toarray(), P64.toarray().astype(np.float32))
    np.testing.assert_array_equal(A.phase_after, O.phase_after(om.astype(np.float32), A.n_mid, A.n_shift, np.float32))
    assert not np.array_equal(A.phase_after, B.phase_after)
