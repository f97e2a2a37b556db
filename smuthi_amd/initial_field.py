"""Plane-wave initial field (initial_field.py:66-328) with the regular SWE
coefficients of all particles computed at once (pwe_to_swe_conversion's
discrete branch, fields/transformations.py:77-130)."""
import numpy as np

from . import fields as flds
from . import special as sf
from .expansions import PlaneWaveExpansion, layer_response


class PlaneWave:
    """Plane wave with smuthi's constructor arguments."""

    def __init__(self, vacuum_wavelength, polar_angle, azimuthal_angle, polarization, amplitude=1,
                 reference_point=None):
        if polarization not in (0, 1):
            raise ValueError('polarization must be 0 (TE) or 1 (TM)')
        if np.isclose(np.cos(polar_angle), 0):
            raise ValueError('propagating waves not defined in the xy-plane')
        self.vacuum_wavelength = vacuum_wavelength
        self.polar_angle = polar_angle
        self.azimuthal_angle = azimuthal_angle
        self.polarization = polarization
        self.amplitude = amplitude
        self.reference_point = list(reference_point) if reference_point else [0, 0, 0]
        self.validity_conditions = []

    def angular_frequency(self):
        return flds.angular_frequency(self.vacuum_wavelength)

    def plane_wave_expansion(self, layer_system, i):
        """(upgoing, downgoing) discrete PWEs in layer i, layer response included."""
        if np.cos(self.polar_angle) > 0:
            iP, kind = 0, 'upgoing'
        else:
            iP, kind = layer_system.number_of_layers() - 1, 'downgoing'
        n = layer_system.refractive_indices[iP]
        om = self.angular_frequency()
        k = n * om
        kx = k * np.sin(self.polar_angle) * np.cos(self.azimuthal_angle)
        ky = k * np.sin(self.polar_angle) * np.sin(self.azimuthal_angle)
        kz = k * np.cos(self.polar_angle)
        z_iP = layer_system.reference_z(iP)
        rp = self.reference_point
        amp = self.amplitude * np.exp(-1j * (kx * rp[0] + ky * rp[1] + kz * (rp[2] - z_iP)))
        pwe = PlaneWaveExpansion(k, np.sin([self.polar_angle]) * n * om, np.array([self.azimuthal_angle]), kind,
                                 [0, 0, z_iP], layer_system.lower_zlimit(iP), layer_system.upper_zlimit(iP))
        pwe.coefficients[self.polarization, 0, 0] = amp
        up, dn = layer_response(layer_system, pwe, iP, i)
        if iP == i:
            if kind == 'upgoing':
                up = up + pwe
            else:
                dn = dn + pwe
        return up, dn

    def swe_coefficients(self, particle_list, layer_system):
        """Regular SWE coefficients a of all particles as one system vector
        (particles grouped by layer and multipole cut-off, one array
        evaluation per group)."""
        sizes = [flds.blocksize(p.l_max, p.m_max) for p in particle_list]
        offsets = np.concatenate([[0], np.cumsum(sizes)])
        out = np.zeros(offsets[-1], complex)
        groups = {}
        for j, p in enumerate(particle_list):
            key = (layer_system.layer_number(p.position[2]), p.l_max, p.m_max)
            groups.setdefault(key, []).append(j)
        for (i, l_max, m_max), idx in groups.items():
            up, dn = self.plane_wave_expansion(layer_system, i)
            pos = np.array([particle_list[j].position for j in idx], dtype=float)
            a = _discrete_pwe_to_swe(up, l_max, m_max, pos) + _discrete_pwe_to_swe(dn, l_max, m_max, pos)
            rows = offsets[np.asarray(idx)][:, None] + np.arange(a.shape[1])[None, :]
            out[rows] = a
        return out

    def swe_coefficients_batched(self, positions, l_max, m_max, layer_system):
        """[n_particles, blocksize] coefficients for particles sharing (l_max, m_max) and one layer."""
        i = layer_system.layer_number(positions[0, 2])
        up, dn = self.plane_wave_expansion(layer_system, i)
        return _discrete_pwe_to_swe(up, l_max, m_max, positions) + _discrete_pwe_to_swe(dn, l_max, m_max, positions)


def _discrete_pwe_to_swe(pwe, l_max, m_max, positions):
    """a[p, n] = 4 sum_pol B-dagger_{n,pol} e^{-i m alpha} g_pol e^{i k.(r_p - r_ref)} for a one-wave PWE."""
    if len(pwe.k_parallel) != 1 or len(pwe.azimuthal_angles) != 1:
        raise ValueError('only discrete plane wave expansions are converted here')
    kp, al = pwe.k_parallel[0], pwe.azimuthal_angles[0]
    kz = pwe.k_z()[0]
    kvec = np.array([kp * np.cos(al), kp * np.sin(al), kz])
    d = positions - np.asarray(pwe.reference_point, dtype=float)[None, :]
    phase = np.exp(1j * (d @ kvec))
    Bd = sf.transformation_arrays(l_max, m_max, np.array([kz / pwe.k]), np.array([kp / pwe.k]), True)[:, :, 0]
    _, _, mm = flds.index_arrays(l_max, m_max)
    g = pwe.coefficients[:, 0, 0]
    per_n = 4 * np.exp(-1j * mm * al) * (Bd[0] * g[0] + Bd[1] * g[1])
    return phase[:, None] * per_n[None, :]
