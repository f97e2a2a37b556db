"""Pins of the oracle's spherical-wave-expansion fields (oracle/swe.py,
restating fields/vector_wave_functions.py:64-155 and fields/expansions.py:348-424)
by closed forms and by the plane-wave representation of the same SWE."""
import numpy as np

from oracle import basics as fb
from oracle import layers as lay
from oracle import swe
from oracle import system as osys


def _points(n, seed, rmin=1.5, rmax=4.0):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return d * rng.uniform(rmin, rmax, n)[:, None]


def test_l1_waves_are_dipole_fields():
    # outgoing tau=1, l=1, m=0 wave = C_E * field of an electric dipole p = z (Jackson 9.18),
    # tau=0, l=1, m=0 = C_M * E of a magnetic dipole m = z; with h_1(x) = -e^{ix}(x + i)/x^2 and the
    # normalized P_1^0 = sqrt(3/2) cos(theta), prefactor 1/2: C_E = -i sqrt(6) / (4 k^3), C_M = -sqrt(6) / (4 k^3)
    for k in (2 * np.pi / 550 * 1.5, 0.012 + 0.0007j):
        p = _points(40, 1) / abs(k)
        r = np.linalg.norm(p, axis=1)
        n = p / r[:, None]
        z = np.array([0.0, 0.0, 1.0])
        e = np.exp(1j * k * r)
        nz = n[:, 2]
        nxz_xn = z[None, :] - n * nz[:, None]
        edip = e[:, None] * (k ** 2 * nxz_xn / r[:, None]
                             + (3 * n * nz[:, None] - z[None, :]) * (1 / r ** 3 - 1j * k / r ** 2)[:, None])
        n_cross_z = np.cross(n, z)
        emag = -k ** 2 * n_cross_z * (e / r * (1 - 1 / (1j * k * r)))[:, None]
        N = np.stack(swe.spherical_vector_wave_function(p[:, 0], p[:, 1], p[:, 2], k, 3, 1, 1, 0), axis=1)
        M = np.stack(swe.spherical_vector_wave_function(p[:, 0], p[:, 1], p[:, 2], k, 3, 0, 1, 0), axis=1)
        ce = -1j * np.sqrt(6) / (4 * k ** 3)
        cm = -np.sqrt(6) / (4 * k ** 3)
        assert np.abs(N - ce * edip).max() / np.abs(N).max() < 1e-12
        assert np.abs(M - cm * emag).max() / np.abs(M).max() < 1e-12


def test_curl_identity_every_multipole():
    # curl N_(tau,l,m) = k N_(1-tau,l,m): ties all (tau, l, m) together and is the relation the
    # magnetic field -i k/omega sum b_(1-tau) N_tau expresses (Faraday with e^{-i omega t})
    k = 0.013 + 0.0004j
    pts = _points(12, 2) / abs(k)
    h = 1e-4 / abs(k)

    def field(tau, l, m, q):
        return np.array(swe.spherical_vector_wave_function(q[:, 0], q[:, 1], q[:, 2], k, 3, tau, l, m))

    for l in range(1, 4):
        for m in range(-l, l + 1):
            for tau in range(2):
                d = []
                for ax in range(3):
                    sh = np.zeros(3)
                    sh[ax] = h
                    d.append((field(tau, l, m, pts + sh) - field(tau, l, m, pts - sh)) / (2 * h))
                # d[ax][comp] = d comp / d ax
                curl = np.array([d[1][2] - d[2][1], d[2][0] - d[0][2], d[0][1] - d[1][0]])
                want = k * field(1 - tau, l, m, pts)
                assert np.abs(curl - want).max() / np.abs(want).max() < 1e-6, (tau, l, m)


def test_swe_matches_its_plane_wave_expansion_above():
    # above the particle the outgoing SWE equals its upgoing PWE (swe_to_pwe_conversion,
    # transformations.py:133-226; PWE field, expansions.py:635-948) up to the kappa/alpha quadrature
    wl = 550.0
    ls = lay.LayerSystem([0, 0], [1.3, 1.3])
    k = 2 * np.pi / wl * 1.3
    l_max = m_max = 3
    rng = np.random.default_rng(5)
    B = fb.blocksize(l_max, m_max)
    coef = rng.standard_normal(B) + 1j * rng.standard_normal(B)
    pos = [10.0, -20.0, 150.0]
    kp = fb.reasonable_Sommerfeld_kpar_contour(wl, layer_refractive_indices=[1.3, 1.3], neff_waypoints=[0, 0.9, 1.3 - 0.05j, 1.7, 5],
                                               neff_resolution=2e-3)
    al = np.linspace(0, 2 * np.pi, 181)
    up, _ = osys.swe_to_pwe_conversion(coef, l_max, m_max, k, pos, kp, al, ls)
    x = np.array([0.0, 120.0, -90.0, 40.0])
    y = np.array([0.0, 60.0, 30.0, -110.0])
    z = pos[2] + np.array([300.0, 350.0, 420.0, 320.0])
    want = np.array(up.electric_field(x, y, z))
    got = np.array(swe.swe_electric_field(coef, l_max, m_max, k, pos, x, y, z))
    assert np.abs(got - want).max() / np.abs(want).max() < 2e-3
    omega = fb.angular_frequency(wl)
    want_h = np.array(up.magnetic_field(x, y, z, wl))
    got_h = np.array(swe.swe_magnetic_field(coef, l_max, m_max, k, pos, x, y, z, omega))
    assert np.abs(got_h - want_h).max() / np.abs(want_h).max() < 2e-3


def test_direct_field_masks():
    k = 0.015
    parts = [(np.ones(fb.blocksize(1, 1), complex), 1, 1, [0.0, 0.0, 100.0], 50.0),
             (np.ones(fb.blocksize(1, 1), complex), 1, 1, [300.0, 0.0, 100.0], 50.0)]
    x = np.array([0.0, 300.0, 150.0, 150.0])
    z = np.array([110.0, 100.0, 100.0, 500.0])
    e = swe.direct_field(parts, k, x, np.zeros(4), z, z_range=(0.0, 400.0))
    one = swe.swe_electric_field(parts[1][0], 1, 1, k, parts[1][3], x[:1], [0.0], z[:1])
    assert abs(e[0][0] - one[0][0]) < 1e-15 * max(1.0, abs(one[0][0]))  # inside particle 0: only particle 1
    assert all(np.abs(c[3]) == 0 for c in e)  # outside the layer's z range
