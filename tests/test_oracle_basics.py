"""Pin the oracle building blocks against the reference's golden literals.

Literals are copied from the reference's own tests (file:line cited per test).
"""
import numpy as np
import pytest

from oracle import basics as fb
from oracle import layers as lay


def test_index_enumeration():
    # tests/unit_tests/test_index_conversion.py:8-37
    for lmax, mmax in ((5, 5), (6, 3), (3, 3), (4, 4), (3, 2)):
        idcs = []
        for tau in range(2):
            for l in range(1, lmax + 1):
                mlim = min(l, mmax)
                for m in range(-mlim, mlim + 1):
                    idcs.append(fb.multi_to_single_index(tau, l, m, lmax, mmax))
        assert idcs == list(range(len(idcs)))
        assert fb.blocksize(lmax, mmax) == len(idcs)
    assert fb.blocksize(3, 3) == 30 and fb.blocksize(4, 4) == 48


def test_legendre_against_prototype():
    # tests/unit_tests/vwf_and_transformations_tests/test_spherical_functions.py:20-60
    lmax = 3
    omega = 2 * 3.14 / 550
    kp = omega * np.array([0.01, 0.2, 0.7, 0.99, 1.2, 2 - 0.5j])
    kz = np.sqrt(omega ** 2 - kp ** 2 + 0j)
    kz[kz.imag < 0] = -kz[kz.imag < 0]
    plm, pilm, taulm = fb.legendre_normalized(kz / omega, kp / omega, lmax)
    np.testing.assert_almost_equal(plm[3, 0, 0], 1.870267465826245)
    np.testing.assert_almost_equal(plm[3, 0, 1], 1.649727250184103)
    np.testing.assert_almost_equal(plm[3, 0, 2], -0.300608757357466)
    np.testing.assert_almost_equal(plm[3, 0, 3], -0.382739631607606)
    np.testing.assert_almost_equal(plm[3, 0, 4], -3.226515147957620j)
    np.testing.assert_almost_equal(plm[3, 0, 5], -25.338383323084539 - 22.141864985871653j)
    np.testing.assert_almost_equal(plm[2, 1, 0], 0.019363948460993)
    np.testing.assert_almost_equal(plm[2, 1, 1], 0.379473319220206)
    np.testing.assert_almost_equal(plm[2, 1, 2], 0.968052168015753)
    np.testing.assert_almost_equal(plm[2, 1, 3], 0.270444009917026)
    np.testing.assert_almost_equal(plm[2, 1, 4], 1.541427909439815j)
    np.testing.assert_almost_equal(plm[2, 1, 5], 3.906499971729346 + 6.239600710712296j)
    np.testing.assert_almost_equal(pilm[2, 1, 0], 1.936394846099318)
    np.testing.assert_almost_equal(pilm[2, 1, 1], 1.897366596101028)
    np.testing.assert_almost_equal(pilm[2, 1, 2], 1.382931668593933)
    np.testing.assert_almost_equal(pilm[2, 1, 3], 0.273175767592955)
    np.testing.assert_almost_equal(pilm[2, 1, 4], 1.284523257866512j)
    np.testing.assert_almost_equal(pilm[2, 1, 5], 1.104282256024128 + 3.395870919362181j)
    np.testing.assert_almost_equal(taulm[3, 2, 0], 0.051227068616724)
    np.testing.assert_almost_equal(taulm[3, 2, 1], 0.963213372000203)
    np.testing.assert_almost_equal(taulm[3, 2, 2], 0.950404683542753)
    np.testing.assert_almost_equal(taulm[3, 2, 3], -2.384713931794872)
    np.testing.assert_almost_equal(taulm[3, 2, 4], -7.131877733107878)
    np.testing.assert_almost_equal(taulm[3, 2, 5], -39.706934218093430 + 42.588889121019569j)



def test_legendre_high_orders_against_closed_forms():
    """Orders up to 16 (direct coupling of l_max = 8 blocks): the P_m^m prefactor against the exact
    rational (2m+1)!!/(2 (2m)!!), and P_l^m against scipy's associated Legendre functions with the
    normalisation sqrt((2l+1)/2 (l-m)!/(l+m)!) and without the Condon-Shortley phase."""
    import math
    from fractions import Fraction
    import scipy.special
    for m in range(0, 17):
        exact = Fraction(1)
        for t in range(1, m + 1):
            exact *= Fraction(2 * t + 1, 2 * t)
        assert abs(fb._prefactor(m) - math.sqrt(exact / 2)) <= 4e-16 * math.sqrt(exact / 2)
    th = np.array([0.3, 1.1, 2.0, 2.9])
    plm = fb.legendre_normalized(np.cos(th), np.sin(th), 16)[0]
    for l in range(17):
        for m in range(l + 1):
            norm = math.sqrt((2 * l + 1) / 2 * math.factorial(l - m) / math.factorial(l + m))
            want = (-1) ** m * norm * scipy.special.lpmv(m, l, np.cos(th))
            np.testing.assert_allclose(plm[l, m].real, want, rtol=1e-11, atol=1e-13)
            assert np.all(plm[l, m].imag == 0)

ZS = np.array([0.01, 2, 5, 2 + 0.1j, 3 - 0.2j, 20 + 20j])


def test_bessel_hankel_against_prototype():
    # test_spherical_functions.py:63-108
    j = fb.spherical_bessel(4, ZS)
    for got, want in zip(j, [1.058196248205502e-11, 0.014079392762915, 0.187017655344890,
                             0.013925330885893 + 0.002550081632129j, 0.055554281414152 - 0.011718427699962j,
                             5.430299683226971e+06 - 3.884383001639664e+06j]):
        np.testing.assert_almost_equal(got, want)
    h = fb.spherical_hankel(4, ZS)
    for got, want in zip(h[1:], [0.014079392762917 - 4.461291526363127j, 0.187017655344889 - 0.186615531479296j,
                                 -0.937540374646528 - 4.322684701489512j, 0.254757423766403 - 0.894658828739464j,
                                 5.349924210583894e-11 - 7.680177127921456e-11j]):
        np.testing.assert_almost_equal(got, want)
    dj = fb.dx_xj(4, ZS)
    for got, want in zip(dj, [5.290971621054867e-11, 0.065126624274088, 0.401032469441925,
                              0.064527362367182 + 0.011261758715092j, 0.230875079050277 - 0.041423344864749j,
                              3.329872586039824e+07 - 1.858505295737451e+08j]):
        np.testing.assert_almost_equal(got, want)
    dh = fb.dx_xh(4, ZS)
    for got, want in zip(dh[1:], [0.065126624274084 + 14.876432990566345j, 0.401032469441923 + 0.669247576352214j,
                                  3.574345443512018 + 14.372976166070977j, -0.423459406818455 + 1.976243655979050j,
                                  4.344741738782320e-10 + 2.612636745427169e-09j]):
        np.testing.assert_almost_equal(got, want)


def test_wigner3j_against_sympy_and_orthogonality():
    # tests/unit_tests/test_wigner3j.py:15-34 compares with sympy
    from sympy.physics.wigner import wigner_3j
    for (l1, l2, ld, m1, m2) in [(1, 1, 2, 1, -1), (3, 2, 3, -1, 2), (4, 3, 5, 3, 1), (3, 3, 6, 3, -3),
                                 (7, 5, 4, 2, -2), (6, 6, 0, 1, 1), (2, 5, 4, -2, 0)]:
        want = float(wigner_3j(l1, l2, ld, m1, -m2, -(m1 - m2)))
        got = fb.wigner3j(l1, l2, ld, m1, -m2, -(m1 - m2))
        assert abs(got - want) < 1e-14
    # orthogonality: sum_{m1,m2} (j1 j2 j3; m1 m2 m3)^2 = 1/(2 j3 + 1)
    j1, j2, j3, m3 = 3, 4, 5, 2
    s = sum(fb.wigner3j(j1, j2, j3, m1, -m3 - m1, m3) ** 2 for m1 in range(-j1, j1 + 1))
    assert abs(s - 1 / (2 * j3 + 1)) < 1e-14


def test_ab5_against_prototype():
    # test_svwf_translation.py:51-59
    a5, _ = fb.ab5_coefficients(3, -1, 2, 2, 3)
    assert abs(a5 - 0.235702260395516) / abs(a5) < 1e-7
    _, b5 = fb.ab5_coefficients(4, 3, 2, 2, 3)
    assert abs(b5 - 0.912870929175277j) / abs(b5) < 1e-7


LAYER_D = [0, 300, 400, 0]
LAYER_N = [1, 2 + 0.1j, 3, 1 + 5j]
OMEGA = 2 * 3.15 / 550
KPAR = OMEGA * 1.7


def test_layer_response_against_tspl_prototype():
    # tests/unit_tests/test_layerresponse.py:43-60
    lmat = lay.layersystem_response_matrix(0, LAYER_D, LAYER_N, KPAR, OMEGA, 2, 1)
    want = np.array([[-0.392979481352895 - 0.376963315605839j, -0.455367266697897 + 0.426065579868901j],
                     [0.545168303416962 - 0.345873455516963j, -0.361796569025878 - 0.644799225334747j]])
    np.testing.assert_almost_equal(lmat, want)
    lmat = lay.layersystem_response_matrix(1, LAYER_D, LAYER_N, KPAR, OMEGA, 1, 2)
    want = np.array([[-0.240373686730040 - 0.148769054113797j, 0.161922209423045 + 0.222085165907288j],
                     [-0.182951011363592 + 0.138158890222525j, 0.215395950986834 - 0.057346289106977j]])
    np.testing.assert_almost_equal(lmat, want)


def test_scattering_matrix_equals_transfer_matrix():
    # test_layerresponse.py:37-41
    for pol in (0, 1):
        t = lay.layersystem_transfer_matrix(pol, LAYER_D, LAYER_N, KPAR, OMEGA)
        s = lay.layersystem_scattering_matrix(pol, LAYER_D, LAYER_N, KPAR, OMEGA)
        np.testing.assert_almost_equal(t[1, 0] / t[0, 0], s[1, 0])


def test_layer_response_mpmath_equals_numpy():
    # test_layerresponse.py:18-35
    import mpmath
    mpmath.mp.dps = 30
    for pol in (0, 1):
        for fl in range(4):
            for tl in range(4):
                a = lay.layersystem_response_matrix(pol, LAYER_D, LAYER_N, KPAR, OMEGA, fl, tl)
                b = lay.layersystem_response_matrix(pol, LAYER_D, LAYER_N, KPAR, OMEGA, fl, tl, mp=mpmath)
                np.testing.assert_almost_equal(a, b)


def test_mie_against_prototype():
    # t_matrix_tests/test_t_matrix.py:23-27
    om = 2 * 3.15 / 550
    q = lay.mie_coefficient(0, 4, om * (1.6 + 0.1j), om * (2.6 + 0.4j), 260)
    np.testing.assert_almost_equal(q, -0.421469104177215 - 0.346698396584972j)
    q = lay.mie_coefficient(1, 4, om * (1.6 + 0.1j), om * (2.6 + 0.4j), 260)
    np.testing.assert_almost_equal(q, -0.511272170262304 - 0.061284547954858j)
    # closed form: index-matched sphere does not scatter
    assert abs(lay.mie_coefficient(0, 2, om * 1.5, om * 1.5, 100)) < 1e-15
    assert abs(lay.mie_coefficient(1, 2, om * 1.5, om * 1.5, 100)) < 1e-15


def test_contour_properties():
    wl = 550
    kp = fb.reasonable_Sommerfeld_kpar_contour(wl, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    neff = kp / fb.angular_frequency(wl)
    # waypoints 0, 0.9, 0.9-0.01j, 1.72-0.01j, 1.72, 2.72 are hit
    for wp in (0, 0.9, 0.9 - 0.01j, 1.72 - 0.01j, 1.72, 2.72):
        assert np.min(abs(neff - wp)) < 1e-12
    steps = abs(np.diff(neff))
    assert steps.max() <= 1e-2 + 1e-12
    assert np.all(abs(neff[:, None] - np.array([1.52, 1])[None, :]) >= 1e-2 / 5)
    # real-axis contour through a branch point gets corrected
    kp2 = fb.reasonable_Sommerfeld_kpar_contour(wl, neff_waypoints=[0, 2], layer_refractive_indices=[1.5, 1],
                                                neff_resolution=0.1)
    n2 = kp2 / fb.angular_frequency(wl)
    assert np.all(abs(n2[:, None] - np.array([1.5, 1])[None, :]) >= 0.1 / 5)
