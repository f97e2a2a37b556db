"""Pin the oracle's whole-solve pipeline and its field evaluation."""
import numpy as np
import pytest

from oracle import basics as fb
from oracle import layers as lay
from oracle import coupling as cp
from oracle import system as osys

WL = 550


def test_single_sphere_on_substrate_against_prototype():
    # tests/system_tests/test_single_sphere_on_substrate.py:14-62
    kp = fb.reasonable_Sommerfeld_kpar_contour(vacuum_wavelength=WL, neff_waypoints=[0, 0.5, 0.8 - 0.1j, 2 - 0.1j,
                                                                                    2.5, 4], neff_resolution=1e-3)
    p = osys.Sphere([0, 0, 150], 2.4, 100, 3, 3)
    ls = lay.LayerSystem([0, 0], [1.52, 1])
    pw = osys.PlaneWave(WL, np.pi, 0, 0, 1, [0, 0, 0])
    lsys = osys.LinearSystemOracle([p], pw, ls, kp)
    b = lsys.solve_lu(lsys.explicit_coupling())
    for idx, want in ((0, -0.2586209 + 0.8111274j), (10, -1.5103858e-04 - 4.1782795e-04j),
                      (21, -0.0795316 + 0.0194518j)):
        assert abs((b[idx] - want) / want) < 1e-4


def _three_spheres():
    kp = fb.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[2, 1.4, 2], neff_resolution=1e-2)
    plist = [osys.Sphere([100, 100, 150], 2.4, 110, 4, 4), osys.Sphere([-100, -100, 250], 1.9, 120, 3, 3),
             osys.Sphere([-200, 100, 300], 1.7, 90, 3, 3)]
    ls = lay.LayerSystem([0, 400, 0], [2, 1.4, 2])
    return kp, plist, ls


def test_gmres_matches_lu():
    # tests/unit_tests/test_gmres.py:13-57 (plane wave instead of the Gaussian beam)
    kp, plist, ls = _three_spheres()
    pw = osys.PlaneWave(WL, np.pi * 7 / 8, np.pi / 3, 0)
    lsys = osys.LinearSystemOracle(plist, pw, ls, kp)
    W = lsys.explicit_coupling()
    b_lu = lsys.solve_lu(W)
    b_gm, info = lsys.solve_gmres(lambda v: W @ v, tol=1e-5)
    assert info == 0
    np.testing.assert_allclose(np.linalg.norm(b_lu), np.linalg.norm(b_gm), rtol=1e-5)
    rhs = lsys.right_hand_side()
    res = np.linalg.norm(rhs - lsys.master_dense(W) @ b_gm) / np.linalg.norm(rhs)
    assert res <= 1e-5


def test_single_sphere_homogeneous_mie_closed_form():
    # one sphere in a homogeneous medium: b = Q a, and sigma_ext = Mie extinction
    kp = fb.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.3, 1.3], neff_resolution=1e-2)
    ls = lay.LayerSystem([0, 0], [1.3, 1.3])
    p = osys.Sphere([0, 0, 100], 1.9 + 0.05j, 120, 6, 6)
    pw = osys.PlaneWave(WL, np.pi, 0, 1)
    lsys = osys.LinearSystemOracle([p], pw, ls, kp)
    W = lsys.explicit_coupling()
    assert np.abs(W).max() == 0
    b = lsys.solve_lu(W)
    np.testing.assert_allclose(b, lsys.t_blocks[0] @ lsys.a, rtol=1e-12, atol=1e-15)
    p.scattered_field = b
    sig = osys.extinction_cross_section(pw, [p], ls)
    k = 2 * np.pi / WL * 1.3
    x = k * 120
    mie = 0.0
    for l in range(1, 7):
        q_te = lay.mie_coefficient(0, l, k, 2 * np.pi / WL * p.refractive_index, 120)
        q_tm = lay.mie_coefficient(1, l, k, 2 * np.pi / WL * p.refractive_index, 120)
        mie += (2 * l + 1) * (q_te + q_tm).real
    mie = -2 * np.pi / k ** 2 * mie
    assert abs(sig - mie) / abs(mie) < 1e-9
    assert x > 1


def test_discrete_pwe_is_plane_wave():
    k = 2 * np.pi / WL * 1.2
    kp, al = 0.6 * k, 0.7
    pwe = osys.PlaneWaveExpansion(k, kp, al, 'upgoing', [10, 20, 30])
    pwe.coefficients[0, 0, 0] = 0.3 + 0.1j
    pwe.coefficients[1, 0, 0] = -0.2 + 0.4j
    x, y, z = np.array([100., -50.]), np.array([30., 200.]), np.array([80., 500.])
    ex, ey, ez = pwe.electric_field(x, y, z)
    kz = np.sqrt(k ** 2 - kp ** 2)
    ph = np.exp(1j * (kp * np.cos(al) * (x - 10) + kp * np.sin(al) * (y - 20) + kz * (z - 30)))
    g0, g1 = pwe.coefficients[:, 0, 0]
    np.testing.assert_allclose(ex, (-np.sin(al) * g0 + np.cos(al) * kz / k * g1) * ph, rtol=1e-13)
    np.testing.assert_allclose(ey, (np.cos(al) * g0 + np.sin(al) * kz / k * g1) * ph, rtol=1e-13)
    np.testing.assert_allclose(ez, (-kp / k * g1) * ph, rtol=1e-13)
    # H = k x E / omega for a plane wave
    hx, hy, hz = pwe.magnetic_field(x, y, z, WL)
    om = 2 * np.pi / WL
    kv = np.array([kp * np.cos(al), kp * np.sin(al), kz])
    E = np.array([ex, ey, ez])
    H = np.cross(kv, E, axis=0) / om
    np.testing.assert_allclose(np.array([hx, hy, hz]), H, rtol=1e-12, atol=1e-16)


def test_pwe_double_trapezoid_brute_force():
    rng = np.random.default_rng(3)
    k = 2 * np.pi / WL
    kpar = np.linspace(0, 1.5, 7) * k + 0.01j * k * np.sin(np.linspace(0, np.pi, 7))
    alpha = np.linspace(0, 2 * np.pi, 9)
    pwe = osys.PlaneWaveExpansion(k, kpar, alpha, 'downgoing', [0, 0, 0])
    pwe.coefficients = rng.standard_normal((2, 7, 9)) + 1j * rng.standard_normal((2, 7, 9))
    x, y, z = np.array([50.0]), np.array([-70.0]), np.array([-40.0])
    ex, ey, ez = pwe.electric_field(x, y, z)
    # explicit loop: sum_k 0.5 dk (I_k + I_{k-1}), I_k = kp_k * sum_a 0.5 da (f_ka + f_k,a-1)
    tot = np.zeros(3, dtype=complex)
    Ik_prev = None
    for ik in range(7):
        kp_ = kpar[ik]
        kz_up = np.sqrt(k ** 2 - kp_ ** 2 + 0j)
        kz_up = kz_up if kz_up.imag >= 0 else -kz_up
        kz = -kz_up  # downgoing
        Ia = np.zeros(3, dtype=complex)
        prev = None
        for ia in range(9):
            a = alpha[ia]
            g0, g1 = pwe.coefficients[:, ik, ia]
            e = np.exp(1j * (kp_ * np.cos(a) * x[0] + kp_ * np.sin(a) * y[0] + kz * z[0]))
            f = np.array([-np.sin(a) * g0 + np.cos(a) * kz / k * g1, np.cos(a) * g0 + np.sin(a) * kz / k * g1,
                          -kp_ / k * g1]) * e
            if prev is not None:
                Ia += 0.5 * (alpha[ia] - alpha[ia - 1]) * (f + prev)
            prev = f
        Ik = Ia * kp_
        if Ik_prev is not None:
            tot += 0.5 * (kpar[ik] - kpar[ik - 1]) * (Ik + Ik_prev)
        Ik_prev = Ik
    np.testing.assert_allclose([ex[0], ey[0], ez[0]], tot, rtol=1e-12)


def test_pwe_to_swe_selection_rule():
    # a plane wave along -z couples only to m = +-1 (closed form of the
    # azimuthal integral); the values themselves are pinned by the
    # single-sphere prototype above, whose RHS is this conversion
    k = 2 * np.pi / WL
    pwe = osys.PlaneWaveExpansion(k, 0.0, 0.0, 'downgoing', [0, 0, 0])
    pwe.coefficients[0, 0, 0] = 1.0
    c = osys.pwe_to_swe_conversion(pwe, 6, 6, [0, 0, 0])
    for n, tau, l, m in fb.index_specs(6, 6):
        if abs(m) != 1:
            assert abs(c[n]) < 1e-14
        else:
            assert abs(c[n]) > 1e-3


def test_scattering_cross_section_mie_closed_form():
    # one sphere in a homogeneous medium: the far-field integral of the scattered intensity
    # (postprocessing/far_field.py:223-515) equals the Mie scattering cross section
    # 2 pi / k^2 sum_l (2l+1) (|Q_te|^2 + |Q_tm|^2) up to the angular quadrature
    kp = fb.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.3, 1.3], neff_resolution=1e-2)
    ls = lay.LayerSystem([0, 0], [1.3, 1.3])
    p = osys.Sphere([0, 0, 100], 1.9 + 0.05j, 120, 6, 6)
    pw = osys.PlaneWave(WL, np.pi, 0, 1)
    lsys = osys.LinearSystemOracle([p], pw, ls, kp)
    p.scattered_field = lsys.solve_lu(lsys.explicit_coupling())
    az, po = np.arange(0, 2 * np.pi + 1e-9, np.pi / 180), np.arange(0, np.pi + 1e-9, np.pi / 360)
    sig = osys.total_scattering_cross_section(pw, [p], ls, po, az)
    k = 2 * np.pi / WL * 1.3
    mie = 0.0
    for l in range(1, 7):
        q_te = lay.mie_coefficient(0, l, k, 2 * np.pi / WL * p.refractive_index, 120)
        q_tm = lay.mie_coefficient(1, l, k, 2 * np.pi / WL * p.refractive_index, 120)
        mie += (2 * l + 1) * (abs(q_te) ** 2 + abs(q_tm) ** 2)
    mie = 2 * np.pi / k ** 2 * mie
    assert abs(sig - mie) / mie < 1e-4


def test_optical_theorem_lossless_waveguide():
    # tests/system_tests/conservation_of_energy_tests/test_conservation_of_energy_planewave.py:14-72:
    # lossless spheres in a lossless layer system, scattering = extinction cross section to 1e-4
    kp = fb.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 5],
                                               neff_resolution=5e-3)
    plist = [osys.Sphere([100, 100, 150], 2.4, 110, 4, 4), osys.Sphere([-100, -100, 250], 1.9, 120, 3, 3),
             osys.Sphere([-200, 100, 300], 1.7, 90, 3, 3)]
    ls = lay.LayerSystem([0, 400, 0], [2, 1.3, 2])
    pw = osys.PlaneWave(WL, np.pi * 7 / 8, np.pi / 3, 0, 1, [0, 0, 400])
    lsys = osys.LinearSystemOracle(plist, pw, ls, kp)
    lsys.scatter_solution(lsys.solve_lu(lsys.explicit_coupling()))
    az, po = np.arange(0, 2 * np.pi + 1e-9, np.pi / 360), np.arange(0, np.pi + 1e-9, np.pi / 360)
    scs = osys.total_scattering_cross_section(pw, plist, ls, po, az)
    ecs = osys.extinction_cross_section(pw, plist, ls)
    assert abs(scs - ecs) / abs(scs) < 1e-4
