"""CPU FP64 restatement of smuthi's non-plane-wave initial fields (test
infrastructure only: the product path never imports this package).

* GaussianBeam.plane_wave_expansion: initial_field.py:157-233 (the beam's
  angular spectrum in the excitation layer, then the layer response);
  its regular SWE at a particle is system.initial_field_coefficients
  (initial_field.py:92-106 over transformations.py:77-130).
* DipoleSource.outgoing_spherical_wave_expansion: initial_field.py:366-388;
  DipoleSource.spherical_wave_expansion: initial_field.py:390-411 (the
  explicit coupling blocks from a virtual l_max = m_max = 1 particle at the
  dipole, applied to that outgoing SWE).
* DipoleCollection.spherical_wave_expansion: initial_field.py:684-709 (sum
  over the dipoles).

Pinned in tests/test_oracle_initial.py by the reference's golden vectors
(tests/unit_tests/initial_field_tests/test_initial_field_gauss.py: SWE
coefficients 0, 9, 20 and |E| = amplitude at the focus;
test_initial_field_dipole.py: SWE coefficients 0, 10, 20, 29 through a
four-layer system), by the large-waist limit (a Gaussian beam tends to the
plane wave of the same direction) and, for the dipole, by the electric
field of its outgoing SWE equal to the textbook dipole field.
"""
import numpy as np

from . import basics as fb
from . import coupling as oc
from . import swe as oswe
from .system import PlaneWaveExpansion, layer_response


class GaussianBeam:
    """Gaussian beam (initial_field.py:157-233)."""

    def __init__(self, vacuum_wavelength, polar_angle, azimuthal_angle, polarization, beam_waist, k_parallel_array,
                 azimuthal_angles_array, amplitude=1, reference_point=None):
        self.vacuum_wavelength = vacuum_wavelength
        self.polar_angle = polar_angle
        self.azimuthal_angle = azimuthal_angle
        self.polarization = polarization
        self.beam_waist = beam_waist
        self.k_parallel_array = np.asarray(k_parallel_array)
        self.azimuthal_angles_array = np.asarray(azimuthal_angles_array)
        self.amplitude = amplitude
        self.reference_point = list(reference_point) if reference_point else [0, 0, 0]

    def plane_wave_expansion(self, layer_system, i):
        if np.cos(self.polar_angle) > 0:
            iG, kind = 0, 'upgoing'
        else:
            iG, kind = layer_system.number_of_layers() - 1, 'downgoing'
        k_iG = layer_system.refractive_indices[iG] * fb.angular_frequency(self.vacuum_wavelength)
        z_iG = layer_system.reference_z(iG)
        pwe = PlaneWaveExpansion(k_iG, self.k_parallel_array, self.azimuthal_angles_array, kind, [0, 0, z_iG],
                                 layer_system.lower_zlimit(iG), layer_system.upper_zlimit(iG))
        kGx = k_iG * np.sin(self.polar_angle) * np.cos(self.azimuthal_angle)
        kGy = k_iG * np.sin(self.polar_angle) * np.sin(self.azimuthal_angle)
        kp, al, kz = pwe.grids()
        kx, ky = kp * np.cos(al), kp * np.sin(al)
        w = self.beam_waist
        r = self.reference_point
        g = (self.amplitude * w ** 2 / (4 * np.pi) * np.exp(-w ** 2 / 4 * ((kx - kGx) ** 2 + (ky - kGy) ** 2))
             * np.exp(-1j * (kx * r[0] + ky * r[1] + kz * (r[2] - z_iG))))
        ang = al - self.azimuthal_angle + self.polarization * np.pi / 2
        pwe.coefficients[0] = g * np.cos(ang)
        pwe.coefficients[1] = (1 if np.cos(self.polar_angle) > 0 else -1) * g * np.sin(ang)
        up, dn = layer_response(layer_system, pwe, iG, i)
        if iG == i:
            if kind == 'upgoing':
                up = up + pwe
            else:
                dn = dn + pwe
        return up, dn


class _Virtual:
    def __init__(self, position, l_max, m_max):
        self.position = list(position)
        self.l_max = l_max
        self.m_max = m_max


class DipoleSource:
    """Point dipole (initial_field.py:331-640)."""

    def __init__(self, vacuum_wavelength, dipole_moment, position, k_parallel_array):
        self.vacuum_wavelength = vacuum_wavelength
        self.dipole_moment = list(dipole_moment)
        self.position = list(position)
        self.k_parallel_array = np.asarray(k_parallel_array)

    def current(self):
        om = fb.angular_frequency(self.vacuum_wavelength)
        return [-1j * om * d for d in self.dipole_moment]

    def outgoing_coefficients(self, layer_system):
        """b[(tau, 1, m)] of the outgoing l_max = m_max = 1 SWE (initial_field.py:366-388)."""
        om = fb.angular_frequency(self.vacuum_wavelength)
        k = layer_system.refractive_indices[layer_system.layer_number(self.position[2])] * om
        j = self.current()
        b = np.zeros(fb.blocksize(1, 1), dtype=complex)
        for tau in range(2):
            for m in range(-1, 2):
                e = [complex(np.squeeze(c)) for c in
                     oswe.spherical_vector_wave_function(np.zeros(1), np.zeros(1), np.zeros(1), k, 1, tau, 1, -m)]
                b[fb.multi_to_single_index(tau, 1, m, 1, 1)] = (1j * k / np.pi * 1j * om
                                                               * (e[0] * j[0] + e[1] * j[1] + e[2] * j[2]))
        return b

    def spherical_wave_expansion(self, particle, layer_system):
        virtual = _Virtual(self.position, 1, 1)
        wd = oc.direct_coupling_block(self.vacuum_wavelength, particle, virtual, layer_system)
        wr = oc.layer_mediated_coupling_block(self.vacuum_wavelength, particle, virtual, layer_system,
                                              self.k_parallel_array)
        return (wd + wr) @ self.outgoing_coefficients(layer_system)


class DipoleCollection:
    """Sum of dipole sources (initial_field.py:642-709, the per-dipole branch)."""

    def __init__(self, dipoles):
        self.dipole_list = list(dipoles)

    def spherical_wave_expansion(self, particle, layer_system):
        return sum(d.spherical_wave_expansion(particle, layer_system) for d in self.dipole_list)
