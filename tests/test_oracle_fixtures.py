"""The oracle against the reference's system-test literals.

* A dipole-illuminated grid of 21 spheres (l_max 4..7, one medium), explicit
  coupling and LU (tests/system_tests/test_direct_radial_and_volume_particle_coupling.py:26-84):
  all 1748 scattered-field coefficients of the '2D' and '3D' layouts, held in
  tests/golden/dipole_grid.npz, under the reference's own check (relative
  1e-7 on the value, its real and its imaginary part, absolute 1e-18).
* Two spheres in a slab waveguide under a plane wave
  (tests/system_tests/test_two_spheres_in_waveguide.py:14-82): four scattered
  coefficients to 1e-4 and the scattered power in the top and bottom
  half-spaces to 1e-3, the reference's tolerances.
"""
import os

import numpy as np
import scipy.linalg

from oracle import basics as ob
from oracle import coupling as oc
from oracle import initial as oin
from oracle import layers as ol
from oracle import system as osys

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def is_same(want, got):
    """The reference's element check (test_direct_radial_and_volume_particle_coupling.py:20-23)."""
    for a, b in ((want, got), (want.real, got.real), (want.imag, got.imag)):
        if not abs(a - b) <= max(1e-7 * max(abs(a), abs(b)), 1e-18):
            return False
    return True


def dipole_grid(dimension):
    """(spheres, dipole position) of the reference's grid: a centre sphere and four rings of five at
    radii 2 sqrt(2) 120 nm x (1, 2, 3, 4), ring idx with l_max 4 + idx, at z = 0 ('2D') or z = idx r
    ('3D'); all r = 120 nm, n = 5 + 0.5i; dipole (1, 1, 1) at z = 70 nm."""
    radius, l_max = 120, 4
    thetas = np.linspace(0, 1.5 * np.pi, num=5, endpoint=False)
    rs = 2 * np.sqrt(2) * radius * np.linspace(1, 4, num=4, endpoint=True)
    spec = [([0, 0, 0], l_max)]
    for idx, r in enumerate(rs):
        for th in thetas:
            spec.append(([r * np.cos(th), r * np.sin(th), 0 if dimension == '2D' else idx * r], l_max + idx))
    return spec, [0, 0, radius - 50]


def _dipole_grid_solution(dimension):
    wl = 416
    spec, dpos = dipole_grid(dimension)
    parts = [osys.Sphere(p, 5 + 0.5j, 120, l) for p, l in spec]
    ls = ol.LayerSystem([0, 0], [1, 1])
    dip = oin.DipoleSource(wl, [1, 1, 1], dpos, np.array([0.0]))
    _, _, _, offs = oc.system_index_maps(parts)
    n = int(offs[-1])
    W = np.zeros((n, n), complex)
    for i, p in enumerate(parts):
        for j, q in enumerate(parts):
            W[offs[i]:offs[i + 1], offs[j]:offs[j + 1]] = oc.direct_coupling_block(wl, p, q, ls)
    a = np.concatenate([dip.spherical_wave_expansion(p, ls) for p in parts])
    T = scipy.linalg.block_diag(*[p.t_matrix(wl, 1.0) for p in parts])
    return np.linalg.solve(np.eye(n) - T @ W, T @ a)


def test_oracle_reproduces_dipole_grid_golden():
    golden = np.load(os.path.join(GOLDEN, 'dipole_grid.npz'))
    for dim in ('2D', '3D'):
        want = golden['known_%s' % dim.lower()]
        got = _dipole_grid_solution(dim)
        assert got.shape == want.shape == (1748,)
        bad = [i for i in range(len(want)) if not is_same(want[i], got[i])]
        assert not bad, (dim, bad[:5])


WG_WL = 550
WG_LITERALS = {0: -0.8609045 + 0.4019615j, 1: 0.0223519 + 0.0675757j, 27: -0.0258275 + 0.1145384j,
               30: 0.2065701 + 0.1197903j}
WG_TOP, WG_BOTTOM = 4.3895865e+02, 2.9024410e+04


def hemisphere_power(ff, top):
    """Angular integral of the far-field intensity over one hemisphere (FarField.top / bottom and
    integral, far_field.py:142-209)."""
    sel = ff.polar_angles <= np.pi / 2 if top else ff.polar_angles >= np.pi / 2
    beta = ff.polar_angles[sel]
    inner = np.trapezoid(ff.signal[:, sel, :], ff.azimuthal_angles[None, None, :]) * np.sin(beta)[None, :]
    return float(np.sum(np.trapezoid(inner, beta[None, :])))


def test_oracle_reproduces_two_spheres_in_waveguide():
    ls = ol.LayerSystem([0, 400, 0], [1.5, 1.7, 1])
    parts = [osys.Sphere([100, 100, 150], 2.4, 120, 3), osys.Sphere([-100, -100, 250], 1.9 + 0.1j, 120, 3)]
    pw = osys.PlaneWave(WG_WL, np.pi, 0, 0, 1, [0, 0, 400])
    kp = ob.reasonable_Sommerfeld_kpar_contour(WG_WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 20],
                                              neff_resolution=1e-3)
    lso = osys.LinearSystemOracle(parts, pw, ls, kp)
    b = lso.solve_lu(lso.explicit_coupling())
    for k, v in WG_LITERALS.items():
        assert abs((b[k] - v) / v) < 1e-4, (k, b[k])
    lso.scatter_solution(b)
    res = np.pi / 360
    az, pol = np.arange(0, 2 * np.pi + 0.5 * res, res), np.arange(0, np.pi + 0.5 * res, res)
    ff = osys.scattered_far_field(WG_WL, parts, ls, pol, az)
    assert abs(hemisphere_power(ff, True) - WG_TOP) < 1e-3 * WG_TOP
    assert abs(hemisphere_power(ff, False) - WG_BOTTOM) < 1e-3 * WG_BOTTOM
