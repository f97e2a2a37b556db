"""CPU FP64 restatement of smuthi's spherical-wave-expansion fields (test
infrastructure only: the product path never imports this package).

* spherical_vector_wave_function: fields/vector_wave_functions.py:64-155
* swe_electric_field / swe_magnetic_field: fields/expansions.py:348-424
* direct_field: the particles' scattered-field SWEs as summed by
  PiecewiseFieldExpansion (postprocessing/scattered_field.py:72-75,
  fields/expansions.py:151-195): each SWE contributes in its layer's z range
  outside its own particle.

Pinned in tests/test_oracle_swe.py by closed forms: the l = 1 waves equal
the textbook electric / magnetic dipole fields up to the normalization
constant derived in the test, the curl identity curl N_tau = k N_(1-tau)
(which ties every (tau, l, m) together and fixes the magnetic field's form),
and agreement above the particle with the SWE's plane-wave expansion
(swe_to_pwe_conversion + PWE field, both pinned elsewhere).
"""
import numpy as np

from . import basics as sf


def spherical_vector_wave_function(x, y, z, k, nu, tau, l, m):
    """(Ex, Ey, Ez) of the SVWF (nu = 1 regular, 3 outgoing), Doicu-Wriedt-Eremin
    conventions as vector_wave_functions.py:64-155."""
    x, y, z = (np.asarray(a, dtype=float) for a in (x, y, z))
    r = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    non0 = r != 0
    theta = np.zeros(x.shape)
    theta[non0] = np.arccos(z[non0] / r[non0])
    phi = np.arctan2(y, x)
    er = [np.zeros(x.shape), np.zeros(x.shape), np.ones(x.shape)]
    for c, a in enumerate((x, y, z)):
        er[c][non0] = a[non0] / r[non0]
    eth = [np.cos(theta) * np.cos(phi), np.cos(theta) * np.sin(phi), -np.sin(theta)]
    eph = [-np.sin(phi), np.cos(phi), np.zeros(x.shape)]
    plm_l, pilm_l, taulm_l = sf.legendre_normalized(np.cos(theta), np.sin(theta), l)
    plm, pilm, taulm = plm_l[l, abs(m)], pilm_l[l, abs(m)], taulm_l[l, abs(m)]
    kr = k * r
    if nu == 1:
        bes = sf.spherical_bessel(l, kr)
        dxxz = sf.dx_xj(l, kr)
        bes_kr = np.zeros(kr.shape, dtype=complex)
        dxxz_kr = np.zeros(kr.shape, dtype=complex)
        zero = r == 0
        bes_kr[zero] = (l == 1) / 3
        dxxz_kr[zero] = (l == 1) * 2 / 3
        bes_kr[~zero] = bes[~zero] / kr[~zero]
        dxxz_kr[~zero] = dxxz[~zero] / kr[~zero]
    elif nu == 3:
        bes = sf.spherical_hankel(l, kr)
        dxxz = sf.dx_xh(l, kr)
        bes_kr = bes / kr
        dxxz_kr = dxxz / kr
    else:
        raise ValueError('nu must be 1 or 3')
    eimphi = np.exp(1j * m * phi)
    prefac = 1 / np.sqrt(2 * l * (l + 1))
    if tau == 0:
        return tuple(prefac * bes * (1j * m * pilm * eth[c] - taulm * eph[c]) * eimphi for c in range(3))
    return tuple(prefac * (l * (l + 1) * bes_kr * plm * er[c] + dxxz_kr * (taulm * eth[c] + 1j * m * pilm * eph[c]))
                 * eimphi for c in range(3))


def swe_electric_field(coefficients, l_max, m_max, k, reference_point, x, y, z, kind='outgoing'):
    """E of one SWE at points (expansions.py:348-381, without the validity mask)."""
    xr, yr, zr = (np.asarray(a, dtype=float) - reference_point[c] for c, a in enumerate((x, y, z)))
    nu = 3 if kind == 'outgoing' else 1
    out = [np.zeros(xr.shape, dtype=complex) for _ in range(3)]
    for tau in range(2):
        for m in range(-m_max, m_max + 1):
            for l in range(max(1, abs(m)), l_max + 1):
                b = coefficients[sf.multi_to_single_index(tau, l, m, l_max, m_max)]
                N = spherical_vector_wave_function(xr, yr, zr, k, nu, tau, l, m)
                for c in range(3):
                    out[c] += b * N[c]
    return tuple(out)


def swe_magnetic_field(coefficients, l_max, m_max, k, reference_point, x, y, z, omega, kind='outgoing'):
    """H of one SWE (expansions.py:383-424): -i k / omega sum_n b_(1-tau, l, m) N_(tau, l, m)."""
    xr, yr, zr = (np.asarray(a, dtype=float) - reference_point[c] for c, a in enumerate((x, y, z)))
    nu = 3 if kind == 'outgoing' else 1
    out = [np.zeros(xr.shape, dtype=complex) for _ in range(3)]
    for tau in range(2):
        for m in range(-m_max, m_max + 1):
            for l in range(max(1, abs(m)), l_max + 1):
                b = coefficients[sf.multi_to_single_index(1 - tau, l, m, l_max, m_max)]
                N = spherical_vector_wave_function(xr, yr, zr, k, nu, tau, l, m)
                for c in range(3):
                    out[c] += b * N[c]
    return tuple(-1j * k / omega * o for o in out)


def direct_field(particles, k, x, y, z, field='E', omega=1.0, z_range=(-np.inf, np.inf)):
    """Sum of the particles' outgoing SWEs; particles: list of (coefficients,
    l_max, m_max, position, radius). A particle's SWE contributes at points
    with z in z_range and outside its sphere (is_outside, particles.py:122-124)."""
    x, y, z = (np.atleast_1d(np.asarray(a, dtype=float)) for a in (x, y, z))
    out = [np.zeros(x.shape, dtype=complex) for _ in range(3)]
    inlayer = (z >= z_range[0]) & (z < z_range[1])
    for coef, l_max, m_max, pos, radius in particles:
        vld = inlayer & ((x - pos[0]) ** 2 + (y - pos[1]) ** 2 + (z - pos[2]) ** 2 > radius ** 2)
        if not vld.any():
            continue
        if field == 'E':
            f = swe_electric_field(coef, l_max, m_max, k, pos, x[vld], y[vld], z[vld])
        else:
            f = swe_magnetic_field(coef, l_max, m_max, k, pos, x[vld], y[vld], z[vld], omega)
        for c in range(3):
            out[c][vld] += f[c]
    return tuple(out)
