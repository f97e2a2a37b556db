"""SVWF indexing, wavenumbers and Sommerfeld contours (smuthi.fields,
fields/__init__.py:12-370), plus the module-global default arrays smuthi keeps
there. Host-side setup; nothing here runs per matvec.
"""
import numpy as np

default_azimuthal_angles = None
default_polar_angles = None
default_Sommerfeld_k_parallel_array = None
default_initial_field_k_parallel_array = None


def angular_frequency(vacuum_wavelength):
    return 2 * np.pi / vacuum_wavelength


def multi_to_single_index(tau, l, m, l_max, m_max):
    """Single index of (tau, l, m): tau-major, then l, then m."""
    per_tau = m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1)
    before_l = (l - 1) * (l + 1) if l - 1 <= m_max else m_max * (m_max + 2) + (l - 1 - m_max) * (2 * m_max + 1)
    return tau * per_tau + before_l + m + min(l, m_max)


def blocksize(l_max, m_max):
    return 2 * (m_max * (m_max + 2) + (l_max - m_max) * (2 * m_max + 1))


def index_arrays(l_max, m_max):
    """(tau, l, m) arrays over the single index n = 0..blocksize-1."""
    tau, ll, mm = [], [], []
    for t in range(2):
        for l in range(1, l_max + 1):
            ml = min(l, m_max)
            for m in range(-ml, ml + 1):
                tau.append(t)
                ll.append(l)
                mm.append(m)
    return np.array(tau), np.array(ll), np.array(mm)


def k_z(k_parallel=None, n_effective=None, k=None, omega=None, vacuum_wavelength=None, refractive_index=None):
    """sqrt(k^2 - kappa^2) on the branch with non-negative imaginary part."""
    if k_parallel is None:
        omega = angular_frequency(vacuum_wavelength) if omega is None else omega
        k_parallel = n_effective * omega
    if k is None:
        omega = angular_frequency(vacuum_wavelength) if omega is None else omega
        k = refractive_index * omega
    kz = np.sqrt(k ** 2 - np.asarray(k_parallel) ** 2 + 0j)
    return np.where(kz.imag < 0, -kz, kz)


def angular_arrays(angular_resolution=np.pi / 360):
    az = np.arange(0, 2 * np.pi + 0.5 * angular_resolution, angular_resolution, dtype=float)
    po = np.arange(0, np.pi + 0.5 * angular_resolution, angular_resolution, dtype=float)
    return az, po


def reasonable_neff_waypoints(layer_refractive_indices=None, neff_imag=1e-2, neff_max=None, neff_max_offset=1):
    n = np.array([1, 2] if layer_refractive_indices is None else layer_refractive_indices)
    lo = max(0, min(n.real) - 0.1)
    hi = max(n.real) + 0.2
    if neff_max is None:
        neff_max = hi + neff_max_offset
    if neff_imag == 0:
        return [0, neff_max]
    wp = [0, lo, lo - 1j * neff_imag, hi - 1j * neff_imag, hi]
    if neff_max > hi:
        wp.append(neff_max)
    return wp


def create_neff_array(neff_waypoints, neff_resolution):
    """Segments between waypoints, each including both end points."""
    segs = []
    for a, b in zip(neff_waypoints[:-1], neff_waypoints[1:]):
        d = abs(b - a)
        if d > 0:
            t = np.linspace(0, 1, int(np.ceil(d / neff_resolution)) + 1, dtype=complex)
            segs.append(a + t * (b - a))
    return np.concatenate(segs)


def branchpoint_correction(layer_refractive_indices, neff, min_distance):
    """Replace samples closer than min_distance to a branch point by midpoints."""
    for n in layer_refractive_indices:
        while True:
            hits = np.flatnonzero(abs(neff - n) < min_distance)
            if hits.size == 0:
                break
            i = hits[0]
            if i != len(neff) - 1:
                neff = np.insert(neff, i + 1, (neff[i] + neff[i + 1]) / 2.0)
                if abs(neff[i + 1] - n) < min_distance:
                    neff = np.delete(neff, i + 1)
            if i != 0:
                neff[i] = (neff[i - 1] + neff[i]) / 2.0
                if abs(neff[i] - n) < min_distance:
                    neff = np.delete(neff, i)
    return neff


def reasonable_Sommerfeld_neff_contour(neff_waypoints=None, layer_refractive_indices=None, neff_imag=1e-2,
                                       neff_max=None, neff_max_offset=1, neff_resolution=1e-2,
                                       neff_minimal_branchpoint_distance=None):
    if neff_waypoints is None:
        neff_waypoints = reasonable_neff_waypoints(layer_refractive_indices, neff_imag, neff_max, neff_max_offset)
    if neff_minimal_branchpoint_distance is None:
        neff_minimal_branchpoint_distance = neff_resolution / 5
    neff = create_neff_array(neff_waypoints, neff_resolution)
    if layer_refractive_indices is not None:
        neff = branchpoint_correction(layer_refractive_indices, neff, neff_minimal_branchpoint_distance)
    return neff


def reasonable_Sommerfeld_kpar_contour(vacuum_wavelength, neff_waypoints=None, layer_refractive_indices=None,
                                       neff_imag=1e-2, neff_max=None, neff_max_offset=1, neff_resolution=1e-2,
                                       neff_minimal_branchpoint_distance=None):
    return reasonable_Sommerfeld_neff_contour(neff_waypoints, layer_refractive_indices, neff_imag, neff_max,
                                              neff_max_offset, neff_resolution,
                                              neff_minimal_branchpoint_distance) * angular_frequency(vacuum_wavelength)
