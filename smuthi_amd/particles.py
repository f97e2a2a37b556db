"""Spheres and their Mie T-matrices (particles.py:97-147, t_matrix.py:8-93)."""
from functools import lru_cache

import numpy as np
import scipy.special

from . import fields as flds


class Sphere:
    """Sphere particle with smuthi's constructor arguments."""

    def __init__(self, position=None, refractive_index=1 + 0j, radius=1, l_max=None, m_max=None):
        self.position = [0.0, 0.0, 0.0] if position is None else [float(c) for c in position]
        self.refractive_index = refractive_index
        self.radius = radius
        self.l_max = l_max
        self.m_max = l_max if m_max is None else m_max
        self.euler_angles = [0, 0, 0]
        self.initial_field = None
        self.scattered_field = None
        self.t_matrix = None

    def circumscribing_sphere_radius(self):
        return self.radius

    def is_inside(self, x, y, z):
        return (x - self.position[0]) ** 2 + (y - self.position[1]) ** 2 + (z - self.position[2]) ** 2 <= \
            self.radius ** 2

    def is_outside(self, x, y, z):
        return (x - self.position[0]) ** 2 + (y - self.position[1]) ** 2 + (z - self.position[2]) ** 2 > \
            self.radius ** 2

    def compute_t_matrix(self, vacuum_wavelength, n_medium):
        return np.diag(mie_diagonal(self.radius, self.refractive_index, self.l_max, self.m_max, vacuum_wavelength,
                                    n_medium))


def _riccati(l, z):
    """j_l(z), d(z j_l)/dz, h_l(z), d(z h_l)/dz for complex or real z."""
    j = scipy.special.spherical_jn(l, z)
    jm = scipy.special.spherical_jn(l - 1, z)
    y = scipy.special.spherical_yn(l, z)
    ym = scipy.special.spherical_yn(l - 1, z)
    h, hm = j + 1j * y, jm + 1j * ym
    return j, z * jm - l * j, h, z * hm - l * h


def mie_coefficient(tau, l, k_medium, k_particle, radius):
    """Q_{tau l}: tau 0 (spherical TE) or 1 (spherical TM)."""
    jm, djm, hm, dhm = _riccati(l, k_medium * radius)
    jp, djp, _, _ = _riccati(l, k_particle * radius)
    if tau == 0:
        return (jm * djp - jp * djm) / (jp * dhm - hm * djp)
    km2, kp2 = k_medium ** 2, k_particle ** 2
    return (km2 * jm * djp - kp2 * jp * djm) / (kp2 * jp * dhm - km2 * hm * djp)


@lru_cache(maxsize=4096)
def mie_diagonal(radius, refractive_index, l_max, m_max, vacuum_wavelength, n_medium):
    """Diagonal of the sphere T-matrix in multipole-index order."""
    om = flds.angular_frequency(vacuum_wavelength)
    tau, ll, _ = flds.index_arrays(l_max, m_max)
    q = {(t, l): mie_coefficient(t, l, om * n_medium, om * refractive_index, radius)
         for t in range(2) for l in range(1, l_max + 1)}
    return np.array([q[(t, l)] for t, l in zip(tau, ll)], dtype=complex)
