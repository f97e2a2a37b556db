"""Pins of the oracle's non-plane-wave initial fields (oracle/initial.py:
GaussianBeam, initial_field.py:157-233; DipoleSource / DipoleCollection,
initial_field.py:331-709) by the reference's own golden vectors and by
limits and identities the physics fixes."""
import numpy as np

from oracle import basics as fb
from oracle import initial as oin
from oracle import layers as ol
from oracle import system as osys


def _gauss_case():
    # tests/unit_tests/initial_field_tests/test_initial_field_gauss.py:7-23
    ld = 532
    kp = np.linspace(0, 0.99999, 1000) * fb.angular_frequency(ld)
    al = np.linspace(0, 2 * np.pi, 1000)
    beam = oin.GaussianBeam(ld, 0, 0.2 * np.pi, 1, 4000, kp, al, amplitude=1, reference_point=[-100, 100, 200])
    return beam, ol.LayerSystem([0, 0], [1, 1])


def test_gaussian_beam_swe_golden_vector():
    beam, ls = _gauss_case()
    a = osys.initial_field_coefficients(beam, osys.Sphere([100, -200, 300], 1.5, 100, 3, 3), ls)
    # test_initial_field_gauss.py:26-35
    for idx, want in ((0, 0.4040275 - 1.6689055j), (9, -0.0115840 + 0.0107125j),
                      (20, -3.1855342e-04 - 7.7089434e-04j)):
        assert abs((a[idx] - want) / want) < 1e-5, idx


def test_gaussian_beam_focus_amplitude():
    # test_initial_field_gauss.py:38-41: |E| at the focus equals the amplitude
    beam, ls = _gauss_case()
    up, dn = beam.plane_wave_expansion(ls, 1)
    e = [u + d for u, d in zip(up.electric_field([-100], [100], [200]), dn.electric_field([-100], [100], [200]))]
    assert abs(np.sqrt(e[0] ** 2 + e[1] ** 2 + e[2] ** 2)[0] - 1) < 1e-3


def test_gaussian_beam_wide_waist_is_the_plane_wave():
    # w -> infinity: the angular spectrum w^2/(4 pi) e^{-w^2 |k - kG|^2 / 4} has unit weight and
    # shrinks to the beam direction, so near the focus the beam is the plane wave (corrections ~ (lambda/w)^2).
    # For a beam from the top the reference negates the TM spectrum (initial_field.py:222-226) while the
    # plane wave's TM amplitude multiplies the same vector, so there the beam is minus the plane wave.
    ld, w = 550.0, 40000.0
    ls = ol.LayerSystem([0, 200, 0], [1.5, 1.8 + 0.01j, 1.0])
    kp = np.linspace(0, 12 / w, 300)
    al = np.linspace(0, 2 * np.pi, 181)
    for polar, z in ((0.0, -300.0), (np.pi, 320.0)):
        s = osys.Sphere([60, -40, z], 1.5, 80, 2, 2)
        for pol in (0, 1):
            beam = oin.GaussianBeam(ld, polar, 0.3, pol, w, kp, al, reference_point=[0, 0, 0])
            a_beam = osys.initial_field_coefficients(beam, s, ls)
            a_wave = osys.initial_field_coefficients(osys.PlaneWave(ld, polar, 0.3, pol), s, ls)
            sign = -1 if (polar > 1 and pol == 1) else 1
            assert np.abs(a_beam - sign * a_wave).max() < 2e-3 * np.abs(a_wave).max(), (polar, pol)


def _dipole_case():
    # tests/unit_tests/initial_field_tests/test_initial_field_dipole.py:8-31
    ld = 550
    kp = fb.reasonable_Sommerfeld_kpar_contour(vacuum_wavelength=ld,
                                               neff_waypoints=[0, 0.8, 0.8 - 0.1j, 2.1 - 0.1j, 2.1, 4],
                                               neff_resolution=1e-2)
    dip = oin.DipoleSource(ld, [1e7, 2e7, 3e7], [300, -200, 100], kp)
    ls = ol.LayerSystem([0, 200, 200, 0], [1, 1.5, 2 + 1e-2j, 1 + 5j])
    return dip, ls


def test_dipole_swe_golden_vector():
    dip, ls = _dipole_case()
    a = dip.spherical_wave_expansion(osys.Sphere([100, 200, 300], 1.5, 100, 3, 3), ls)
    # test_initial_field_dipole.py:34-52
    for idx, want in ((0, -0.042563330382109 - 1.073039889335632j), (10, 0.496704638004303 + 0.908744692802429j),
                      (20, 1.209331750869751 - 2.165398836135864j), (29, -0.801326990127563 - 0.706078171730042j)):
        assert abs((a[idx] - want) / want) < 1e-4, idx


def test_dipole_outgoing_swe_is_the_dipole_field():
    # the outgoing l = 1 SWE of a dipole in a homogeneous medium is the field of the current
    # j = -i omega mu: E = i omega mu0 G j, G the free-space dyadic Green's function (Novotny & Hecht 8.55),
    # in smuthi's units (c = mu0 = 1): E = (k^2 p + ... ) written with j directly
    ld = 550.0
    n = 1.3
    ls = ol.LayerSystem([0, 0], [n, n])
    dip = oin.DipoleSource(ld, [0.3e7, -1.1e7, 2e7], [10.0, 20.0, 30.0], np.array([0.0, 1.0]))
    b = dip.outgoing_coefficients(ls)
    om = fb.angular_frequency(ld)
    k = n * om
    from oracle import swe as oswe
    rng = np.random.default_rng(2)
    pts = rng.standard_normal((20, 3)) * 400 + np.array(dip.position)
    e = np.array(oswe.swe_electric_field(b, 1, 1, k, dip.position, pts[:, 0], pts[:, 1], pts[:, 2]))
    j = np.array(dip.current())
    R = pts - np.array(dip.position)
    r = np.linalg.norm(R, axis=1)
    u = R / r[:, None]
    g = np.exp(1j * k * r) / (4 * np.pi * r)
    uj = u @ j
    a1 = 1 + (1j * k * r - 1) / (k * r) ** 2
    a2 = (3 - 3j * k * r - (k * r) ** 2) / (k * r) ** 2
    want = 1j * om * g[:, None] * (a1[:, None] * j[None, :] + a2[:, None] * uj[:, None] * u)
    assert np.abs(e.T - want).max() < 1e-10 * np.abs(want).max()
