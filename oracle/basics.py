"""Oracle building blocks: SVWF indexing, Sommerfeld contours, Legendre and
Bessel functions, Wigner-3j, translation and transformation coefficients.

Pinning (tests/test_oracle_basics.py):
  * multi_to_single_index / blocksize: the enumeration identity of
    tests/unit_tests/test_index_conversion.py:8-37.
  * legendre_normalized: the MATLAB-prototype literals of
    tests/unit_tests/vwf_and_transformations_tests/test_spherical_functions.py:20-60.
  * spherical_bessel/hankel, dx_xj, dx_xh: literals of the same file :63-132.
  * wigner3j: sympy.physics.wigner.wigner_3j (third-party, as in
    tests/unit_tests/test_wigner3j.py:15-34) plus orthogonality.
  * ab5_coefficients: literals of test_svwf_translation.py:51-59.
  * transformation_coefficients_vwf: the SWE->PWE->SWE round trip of a plane
    wave (closed form) in tests/test_oracle_fields.py.
  * contour functions: waypoint/segment identities (endpoints, spacing,
    branch-point distance) in tests/test_oracle_basics.py.
"""

from fractions import Fraction
import math

import numpy as np
import scipy.special


def angular_frequency(vacuum_wavelength):
    """omega = 2 pi / lambda (fields/__init__.py:12-22)."""
    return 2 * np.pi / vacuum_wavelength


def multi_to_single_index(tau, l, m, l_max, m_max):
    """tau-major, then l, then m single index (fields/__init__.py:29-80)."""
    tau_blocksize = m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1)
    n = tau * tau_blocksize
    if (l - 1) <= m_max:
        n += (l - 1) * (l - 1 + 2)
    else:
        n += m_max * (m_max + 2) + (l - 1 - m_max) * (2 * m_max + 1)
    n += m + min(l, m_max)
    return n


def blocksize(l_max, m_max):
    """Number of coefficients per particle (fields/__init__.py:83-94)."""
    return multi_to_single_index(1, l_max, m_max, l_max, m_max) + 1


def k_z(k_parallel=None, n_effective=None, k=None, omega=None, vacuum_wavelength=None, refractive_index=None):
    """k_z = sqrt(k^2 - kappa^2) with Im >= 0 (fields/__init__.py:101-129)."""
    if k_parallel is None:
        if omega is None:
            omega = angular_frequency(vacuum_wavelength)
        k_parallel = n_effective * omega
    if k is None:
        if omega is None:
            omega = angular_frequency(vacuum_wavelength)
        k = refractive_index * omega
    kz = np.sqrt(k ** 2 - k_parallel ** 2 + 0j)
    kz = (kz.imag >= 0) * kz + (kz.imag < 0) * (-kz)
    return kz


def angular_arrays(angular_resolution=np.pi / 360):
    """Default azimuthal/polar arrays (fields/__init__.py:132-136)."""
    az = np.arange(0, 2 * np.pi + 0.5 * angular_resolution, angular_resolution, dtype=float)
    po = np.arange(0, np.pi + 0.5 * angular_resolution, angular_resolution, dtype=float)
    return az, po


def reasonable_neff_waypoints(layer_refractive_indices=None, neff_imag=1e-2, neff_max=None, neff_max_offset=1):
    """Contour corner points (fields/__init__.py:153-195)."""
    if layer_refractive_indices is None:
        layer_refractive_indices = [1, 2]
    min_waveguide_neff = max(0, min(np.array(layer_refractive_indices).real) - 0.1)
    max_waveguide_neff = max(np.array(layer_refractive_indices).real) + 0.2
    if neff_max is None:
        neff_max = max_waveguide_neff + neff_max_offset
    if neff_imag == 0:
        return [0, neff_max]
    waypoints = [0, min_waveguide_neff, min_waveguide_neff - 1j * neff_imag,
                 max_waveguide_neff - 1j * neff_imag, max_waveguide_neff]
    if neff_max > max_waveguide_neff:
        waypoints.append(neff_max)
    return waypoints


def create_neff_array(neff_waypoints, neff_resolution):
    """Piecewise-linear contour with duplicated joints (fields/__init__.py:198-219)."""
    segs = []
    for i in range(len(neff_waypoints) - 1):
        d = abs(neff_waypoints[i + 1] - neff_waypoints[i])
        if d > 0:
            num = int(np.ceil(d / neff_resolution)) + 1
            arr = np.linspace(0, 1, num=num, endpoint=True, dtype=complex)
            segs.append(neff_waypoints[i] + arr * (neff_waypoints[i + 1] - neff_waypoints[i]))
    return np.concatenate(segs)


def branchpoint_correction(layer_refractive_indices, n_effective_array, neff_minimal_branchpoint_distance):
    """Move contour samples away from branch points (fields/__init__.py:242-276)."""
    for n in layer_refractive_indices:
        while True:
            idcs = np.where(abs(n_effective_array - n) < neff_minimal_branchpoint_distance)[0]
            if len(idcs) == 0:
                break
            idx = idcs[0]
            if not idx == len(n_effective_array) - 1:
                n_effective_array = np.insert(n_effective_array, idx + 1,
                                              (n_effective_array[idx] + n_effective_array[idx + 1]) / 2.0)
                if abs(n_effective_array[idx + 1] - n) < neff_minimal_branchpoint_distance:
                    n_effective_array = np.delete(n_effective_array, idx + 1)
            if not idx == 0:
                n_effective_array[idx] = (n_effective_array[idx - 1] + n_effective_array[idx]) / 2.0
                if abs(n_effective_array[idx] - n) < neff_minimal_branchpoint_distance:
                    n_effective_array = np.delete(n_effective_array, idx)
    return n_effective_array


def reasonable_Sommerfeld_kpar_contour(vacuum_wavelength, neff_waypoints=None, layer_refractive_indices=None,
                                       neff_imag=1e-2, neff_max=None, neff_max_offset=1, neff_resolution=1e-2,
                                       neff_minimal_branchpoint_distance=None):
    """k_parallel Sommerfeld contour (fields/__init__.py:279-370)."""
    if neff_waypoints is None:
        neff_waypoints = reasonable_neff_waypoints(layer_refractive_indices, neff_imag, neff_max, neff_max_offset)
    if neff_minimal_branchpoint_distance is None:
        neff_minimal_branchpoint_distance = neff_resolution / 5
    neff = create_neff_array(neff_waypoints, neff_resolution)
    if layer_refractive_indices is not None:
        neff = branchpoint_correction(layer_refractive_indices, neff, neff_minimal_branchpoint_distance)
    return neff * angular_frequency(vacuum_wavelength)


def _prefactor(m):
    """sqrt((2m+1)!!/(2 (2m)!!)) by the reference's product form (numba_helpers.py:159-175).

    For m > 10 the reference evaluates the same quantity by its asymptotic
    series in 1/m (numba_helpers.py:177-205), which it truncates where the
    series reaches double precision; the product form is used here for every
    m (the two agree to rounding, checked in tests/test_oracle_basics.py)."""
    res = 1.
    for t in range(2, 2 * m + 2, 2):
        res += res / t
    return (res / 2) ** 0.5


def legendre_normalized(ct, st, lmax):
    """Normalized P_l^m, pi_l^m, tau_l^m for complex (ct, st) (utility/math.py:38-104).

    Returns arrays indexed [l, m, *ct.shape]."""
    ct = np.atleast_1d(np.array(ct, dtype=np.complex128))
    st = np.atleast_1d(np.array(st, dtype=np.complex128))
    shp = ct.shape
    plm = np.zeros((lmax + 1, lmax + 1) + shp, dtype=np.complex128)
    pilm = np.zeros_like(plm)
    taulm = np.zeros_like(plm)
    pprimel0 = np.zeros((lmax + 1,) + shp, dtype=np.complex128)
    plm[0, 0] = np.sqrt(2) / 2
    if lmax >= 1:
        plm[1, 0] = np.sqrt(3 / 2) * ct
        pprimel0[1] = np.sqrt(3) * plm[0, 0]
        taulm[1, 0] = -st * pprimel0[1]
    taulm[0, 0] = -st * pprimel0[0]
    for l in range(1, lmax):
        plm[l + 1, 0] = (1 / (l + 1) * np.sqrt((2 * l + 1) * (2 * l + 3)) * ct * plm[l, 0]
                         - l / (l + 1) * np.sqrt((2 * l + 3) / (2 * l - 1)) * plm[l - 1, 0])
        pprimel0[l + 1] = ((l + 1) * np.sqrt((2 * (l + 1) + 1) / (2 * (l + 1) - 1)) * plm[l, 0]
                           + np.sqrt((2 * (l + 1) + 1) / (2 * (l + 1) - 1)) * ct * pprimel0[l])
        taulm[l + 1, 0] = -st * pprimel0[l + 1]
    for m in range(1, lmax + 1):
        pref = _prefactor(m)
        plm[m, m] = pref * st ** m
        pilm[m, m] = pref * st ** (m - 1)
        taulm[m, m] = m * ct * pilm[m, m]
        for l in range(m, lmax):
            plm[l + 1, m] = (np.sqrt((2 * l + 1) * (2 * l + 3) / ((l + 1 - m) * (l + 1 + m))) * ct * plm[l, m]
                             - np.sqrt((2 * l + 3) * (l - m) * (l + m) / ((2 * l - 1) * (l + 1 - m) * (l + 1 + m)))
                             * plm[l - 1, m])
            pilm[l + 1, m] = (np.sqrt((2 * l + 1) * (2 * l + 3) / (l + 1 - m) / (l + 1 + m)) * ct * pilm[l, m]
                              - np.sqrt((2 * l + 3) * (l - m) * (l + m) / (2 * l - 1) / (l + 1 - m) / (l + 1 + m))
                              * pilm[l - 1, m])
            taulm[l + 1, m] = ((l + 1) * ct * pilm[l + 1, m]
                               - (l + 1 + m) * np.sqrt((2 * (l + 1) + 1) * (l + 1 - m) / (2 * (l + 1) - 1) / (l + 1 + m))
                               * pilm[l, m])
    return plm, pilm, taulm


spherical_bessel = scipy.special.spherical_jn


def spherical_hankel(n, x):
    """h_n^(1) = j_n + i y_n, NaN at x == 0 (utility/math.py:110-115)."""
    j = scipy.special.spherical_jn(n, x)
    y = scipy.special.spherical_yn(n, x)
    if hasattr(x, '__len__'):
        y = np.array(y)
        y[np.asarray(x) == 0] = np.nan
    return j + 1j * y


def dx_xj(n, x):
    """d/dx (x j_n(x)) (utility/math.py:118-129)."""
    return x * spherical_bessel(n - 1, x) - n * spherical_bessel(n, x)


def dx_xh(n, x):
    """d/dx (x h_n(x)) (utility/math.py:132-143)."""
    return x * spherical_hankel(n - 1, x) - n * spherical_hankel(n, x)


def _fact(n):
    return math.factorial(n)


def wigner3j(j1, j2, j3, m1, m2, m3):
    """Wigner-3j symbol by the Racah formula in exact rational arithmetic.

    Replaces the pywigxjpf / sympy call of utility/math.py:14-35 (integer j only)."""
    if m1 + m2 + m3 != 0:
        return 0.0
    if j3 < abs(j1 - j2) or j3 > j1 + j2:
        return 0.0
    if abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    tri = Fraction(_fact(j1 + j2 - j3) * _fact(j1 - j2 + j3) * _fact(-j1 + j2 + j3), _fact(j1 + j2 + j3 + 1))
    pre = tri * _fact(j1 + m1) * _fact(j1 - m1) * _fact(j2 + m2) * _fact(j2 - m2) * _fact(j3 + m3) * _fact(j3 - m3)
    s = Fraction(0)
    kmin = max(0, j2 - j3 - m1, j1 - j3 + m2)
    kmax = min(j1 + j2 - j3, j1 - m1, j2 + m2)
    for k in range(kmin, kmax + 1):
        den = (_fact(k) * _fact(j1 + j2 - j3 - k) * _fact(j1 - m1 - k) * _fact(j2 + m2 - k)
               * _fact(j3 - j2 + m1 + k) * _fact(j3 - j1 - m2 + k))
        s += Fraction((-1) ** k, den)
    sign = (-1) ** (j1 - j2 - m3)
    val = s * s * pre
    return sign * math.sqrt(val) * (1 if s >= 0 else -1)


_AB5 = {}


def ab5_coefficients(l1, m1, l2, m2, p):
    """a5, b5 of the SVWF translation operator (fields/transformations.py:413-443)."""
    key = (l1, m1, l2, m2, p)
    if key in _AB5:
        return _AB5[key]
    jfac = 1.0j ** (abs(m1 - m2) - abs(m1) - abs(m2) + l2 - l1 + p) * (-1) ** (m1 - m2)
    fac1 = np.sqrt((2 * l1 + 1) * (2 * l2 + 1) / (2 * l1 * (l1 + 1) * l2 * (l2 + 1)))
    fac2a = (l1 * (l1 + 1) + l2 * (l2 + 1) - p * (p + 1)) * np.sqrt(2 * p + 1)
    fac2b = np.sqrt((l1 + l2 + 1 + p) * (l1 + l2 + 1 - p) * (p + l1 - l2) * (p - l1 + l2) * (2 * p + 1))
    wig1 = wigner3j(l1, l2, p, m1, -m2, -(m1 - m2))
    wig2a = wigner3j(l1, l2, p, 0, 0, 0)
    wig2b = wigner3j(l1, l2, p - 1, 0, 0, 0) if p >= 1 else 0.0
    a = jfac * fac1 * fac2a * wig1 * wig2a
    b = jfac * fac1 * fac2b * wig1 * wig2b
    _AB5[key] = (a, b)
    return a, b


def transformation_coefficients_vwf(tau, l, m, pol, kp=None, kz=None, pilm_list=None, taulm_list=None, dagger=False):
    """SVWF <-> PVWF transformation coefficient B (fields/transformations.py:18-74)."""
    if pilm_list is None:
        k = np.sqrt(kp ** 2 + kz ** 2)
        _, pilm_list, taulm_list = legendre_normalized(kz / k, kp / k, l)
    if tau == pol:
        sphfun = taulm_list[l, abs(m)]
    else:
        sphfun = m * pilm_list[l, abs(m)]
    if dagger:
        if pol == 0:
            prefac = -1 / (-1j) ** (l + 1) / np.sqrt(2 * l * (l + 1)) * (-1j)
        else:
            prefac = -1 / (-1j) ** (l + 1) / np.sqrt(2 * l * (l + 1)) * 1
    else:
        if pol == 0:
            prefac = -1 / (1j) ** (l + 1) / np.sqrt(2 * l * (l + 1)) * (1j)
        else:
            prefac = -1 / (1j) ** (l + 1) / np.sqrt(2 * l * (l + 1)) * 1
    return prefac * sphfun


def index_specs(l_max, m_max):
    """(n, tau, l, m) for all indices of one particle, in index order."""
    out = []
    for tau in range(2):
        for m in range(-m_max, m_max + 1):
            for l in range(max(1, abs(m)), l_max + 1):
                out.append((multi_to_single_index(tau, l, m, l_max, m_max), tau, l, m))
    out.sort()
    return out
