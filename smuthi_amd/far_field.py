"""Far field and cross sections (postprocessing/far_field.py:9-654) from the
scattered-field plane wave expansions, assembled on the GPU: the far field
in the top / bottom layer is the PWE at kappa = k sin(beta)
(pwe_to_ff_conversion), the scattering cross section its angular integral,
the extinction cross section the PWE at the incident direction. For many
right-hand sides solved on one particle configuration,
scattering_cross_sections / scattered_far_fields evaluate every solution
vector without touching the particles' stored fields."""
import numpy as np

from . import fields as flds
from .scattered_field import scattered_field_pwe


class FarField:
    """Far-field amplitude [2 pol, n_beta, n_alpha] and intensity (signal), as
    smuthi's FarField (far_field.py:9-209)."""

    def __init__(self, polar_angles='default', azimuthal_angles='default', angular_resolution=None,
                 signal_type='intensity', reference_point=None):
        azimuthal_angles, polar_angles = _angular_arrays(azimuthal_angles, polar_angles, angular_resolution)
        self.polar_angles = np.asarray(polar_angles)
        self.azimuthal_angles = np.asarray(azimuthal_angles)
        self.reference_point = reference_point
        self.amplitude = np.full((2, len(self.polar_angles), len(self.azimuthal_angles)), np.nan, dtype=complex)
        self.signal = np.full((2, len(self.polar_angles), len(self.azimuthal_angles)), np.nan)
        self.signal_type = signal_type

    def electric_field_amplitude(self):
        """(A_x, A_y, A_z) from the TE (e_phi) and TM (e_theta) amplitudes."""
        eph_x = -np.sin(self.azimuthal_angles)[None, :]
        eph_y = np.cos(self.azimuthal_angles)[None, :]
        eth_x = np.outer(np.cos(self.polar_angles), np.cos(self.azimuthal_angles))
        eth_y = np.outer(np.cos(self.polar_angles), np.sin(self.azimuthal_angles))
        eth_z = -np.sin(self.polar_angles)[:, None]
        a0, a1 = self.amplitude[0], self.amplitude[1]
        return a0 * eph_x + a1 * eth_x, a0 * eph_y + a1 * eth_y, a1 * eth_z

    def azimuthal_integral_times_sin_beta(self):
        if len(self.azimuthal_angles) > 2:
            return np.trapezoid(self.signal, self.azimuthal_angles[None, None, :]) * np.sin(self.polar_angles[None, :])
        return None

    def azimuthal_integral(self):
        if len(self.azimuthal_angles) > 2:
            return np.trapezoid(self.signal, self.azimuthal_angles[None, None, :])
        return None

    def integral(self):
        """Radiated power per polarization [2]."""
        if len(self.azimuthal_angles) > 2:
            return np.trapezoid(self.azimuthal_integral_times_sin_beta(), self.polar_angles[None, :])
        return None

    def _part(self, sel):
        ff = FarField(self.polar_angles[sel], self.azimuthal_angles, signal_type=self.signal_type)
        ff.signal = self.signal[:, sel, :]
        ff.amplitude = self.amplitude[:, sel, :]
        return ff

    def top(self):
        sel = self.polar_angles <= np.pi / 2
        return self._part(sel) if sel.any() else None

    def bottom(self):
        sel = self.polar_angles >= np.pi / 2
        return self._part(sel) if sel.any() else None

    def alpha_grid(self):
        return np.meshgrid(self.azimuthal_angles, self.polar_angles.real)[0]

    def beta_grid(self):
        return np.meshgrid(self.azimuthal_angles, self.polar_angles.real)[1]

    def append(self, other):
        """Join a far field of a disjoint polar range (far_field.py:188-209)."""
        if not np.array_equal(self.azimuthal_angles, other.azimuthal_angles):
            raise ValueError('azimuthal angles not consistent')
        if self.signal_type != other.signal_type:
            raise ValueError('signal type not consistent')
        if max(self.polar_angles) <= min(other.polar_angles):
            first, second = self, other
        elif min(self.polar_angles) >= max(other.polar_angles):
            first, second = other, self
        else:
            raise ValueError('far fields have overlapping polar angle domains')
        self.polar_angles = np.concatenate((first.polar_angles, second.polar_angles))
        self.signal = np.concatenate((first.signal, second.signal), 1)
        self.amplitude = np.concatenate((first.amplitude, second.amplitude), 1)


def _angular_arrays(azimuthal_angles='default', polar_angles='default', angular_resolution=None):
    if angular_resolution is not None:
        return flds.angular_arrays(angular_resolution)
    if isinstance(polar_angles, str) and polar_angles == 'default':
        polar_angles = flds.default_polar_angles
    if isinstance(azimuthal_angles, str) and azimuthal_angles == 'default':
        azimuthal_angles = flds.default_azimuthal_angles
    if polar_angles is None or azimuthal_angles is None:
        azimuthal_angles, polar_angles = flds.angular_arrays()
    return np.asarray(azimuthal_angles, dtype=float), np.asarray(polar_angles)


def pwe_to_ff_conversion(vacuum_wavelength, plane_wave_expansion):
    """Far field of a plane wave expansion (far_field.py:223-259): amplitude
    -2 pi k kz g, intensity 2 pi^2 k kz^2 |g|^2 / omega at the polar angles
    asin(kappa / k) (upgoing) or pi - asin(kappa / k) (downgoing)."""
    pwe = plane_wave_expansion
    omega = flds.angular_frequency(vacuum_wavelength)
    k, kp = pwe.k, np.asarray(pwe.k_parallel)
    if pwe.kind == 'upgoing':
        polar = np.arcsin(kp / k)
    elif pwe.kind == 'downgoing':
        polar = np.pi - np.arcsin(kp / k)
    else:
        raise ValueError('PWE type not specified')
    if np.any(np.imag(polar)):
        raise ValueError('complex angles are not allowed')
    kz = flds.k_z(k_parallel=kp, k=k)
    ampli = -2 * np.pi * (kz * k)[None, :, None] * pwe.coefficients
    intens = (2 * np.pi ** 2 / omega * (kz ** 2 * k)[None, :, None] * abs(pwe.coefficients) ** 2).real
    idx = np.argsort(np.real(polar))
    ff = FarField(polar_angles=np.real(polar)[idx], azimuthal_angles=pwe.azimuthal_angles,
                  reference_point=pwe.reference_point)
    ff.amplitude = ampli[:, idx, :]
    ff.signal = intens[:, idx, :]
    return ff


def scattered_far_field(vacuum_wavelength, particle_list, layer_system, polar_angles='default',
                        azimuthal_angles='default', angular_resolution=None, reference_point=None,
                        precision='double'):
    """Scattered far field in the top and bottom layers (far_field.py:363-427)."""
    if reference_point is not None:
        raise NotImplementedError('far fields are referred to the layer reference points')
    omega = flds.angular_frequency(vacuum_wavelength)
    azimuthal_angles, polar_angles = _angular_arrays(azimuthal_angles, polar_angles, angular_resolution)
    if np.any(np.imag(polar_angles)):
        raise ValueError('complex angles not allowed in far field')
    polar_angles = np.real(polar_angles)
    i_top = layer_system.number_of_layers() - 1
    top = polar_angles[polar_angles < np.pi / 2]
    bot = polar_angles[polar_angles > np.pi / 2]
    n_top = np.sort(np.sin(top) * layer_system.refractive_indices[i_top])
    n_bot = np.sort(np.sin(bot) * layer_system.refractive_indices[0])
    ff_top = ff_bot = None
    if len(top) > 1 and np.imag(layer_system.refractive_indices[i_top]) == 0:
        up, _ = scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, i_top, n_top * omega,
                                    azimuthal_angles, precision=precision)
        ff_top = pwe_to_ff_conversion(vacuum_wavelength, up)
    if len(bot) > 1 and np.imag(layer_system.refractive_indices[0]) == 0:
        _, dn = scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, 0, n_bot * omega,
                                    azimuthal_angles, precision=precision)
        ff_bot = pwe_to_ff_conversion(vacuum_wavelength, dn)
    if ff_top is not None:
        if ff_bot is not None:
            ff_top.append(ff_bot)
        far_field = ff_top
    else:
        far_field = ff_bot
    far_field.polar_angles = far_field.polar_angles.real
    return far_field


def _initial_intensity(initial_field, layer_system):
    i_top = layer_system.number_of_layers() - 1
    n_p = layer_system.refractive_indices[0 if np.cos(initial_field.polar_angle) > 0 else i_top]
    if np.imag(n_p):
        raise ValueError('plane wave from absorbing layer: cross section undefined')
    return abs(initial_field.amplitude) ** 2 * np.real(n_p) / 2


def scattering_cross_section(initial_field, particle_list, layer_system, polar_angles='default',
                             azimuthal_angles='default', angular_resolution=None, precision='double'):
    """Differential scattering cross section as a FarField (far_field.py:429-478)."""
    if type(initial_field).__name__ != 'PlaneWave':
        raise ValueError('Cross section only defined for plane wave excitation.')
    i0 = _initial_intensity(initial_field, layer_system)
    dscs = scattered_far_field(initial_field.vacuum_wavelength, particle_list, layer_system, polar_angles,
                               azimuthal_angles, angular_resolution, precision=precision)
    dscs.signal_type = 'differential scattering cross section'
    dscs.signal = dscs.signal / i0
    return dscs


def total_scattering_cross_section(simulation=None, initial_field=None, particle_list=None, layer_system=None,
                                   polar_angles='default', azimuthal_angles='default', angular_resolution=None,
                                   precision='double'):
    """Angle-integrated scattering cross section (far_field.py:480-515)."""
    if initial_field is None:
        initial_field = simulation.initial_field
    if particle_list is None:
        particle_list = simulation.particle_list
    if layer_system is None:
        layer_system = simulation.layer_system
    scs = scattering_cross_section(initial_field, particle_list, layer_system, polar_angles, azimuthal_angles,
                                   angular_resolution, precision).integral()
    return scs[0] + scs[1]


class _Scatterer:
    """A particle's geometry with the scattered-field coefficients of one solution vector."""

    def __init__(self, particle, coefficients):
        self.position, self.radius = particle.position, getattr(particle, 'radius', None)
        self.l_max, self.m_max = particle.l_max, particle.m_max
        self.scattered_field = coefficients
        self.is_outside = particle.is_outside


def _per_particle(particle_list, x):
    out, off = [], 0
    for p in particle_list:
        n = flds.blocksize(p.l_max, p.m_max)
        out.append(_Scatterer(p, np.asarray(x[off:off + n], dtype=complex)))
        off += n
    if off != len(x):
        raise ValueError('solution length %d does not match the particles (%d unknowns)' % (len(x), off))
    return out


def scattered_far_fields(vacuum_wavelength, particle_list, layer_system, solutions, polar_angles='default',
                         azimuthal_angles='default', angular_resolution=None, precision='double'):
    """Scattered far field of every solution vector ([n_rhs, n_unknowns], e.g.
    LinearSystem.solve of a block of right-hand sides) on one particle configuration."""
    return [scattered_far_field(vacuum_wavelength, _per_particle(particle_list, x), layer_system, polar_angles,
                                azimuthal_angles, angular_resolution, precision=precision)
            for x in np.atleast_2d(solutions)]


def scattering_cross_sections(initial_fields, particle_list, layer_system, solutions, polar_angles='default',
                              azimuthal_angles='default', angular_resolution=None, precision='double'):
    """Total scattering and extinction cross sections [n_rhs] of plane-wave
    right-hand sides initial_fields[j] with solution vectors solutions[j]."""
    scs, ecs = [], []
    for pw, x in zip(initial_fields, np.atleast_2d(solutions)):
        parts = _per_particle(particle_list, x)
        scs.append(total_scattering_cross_section(initial_field=pw, particle_list=parts, layer_system=layer_system,
                                                  polar_angles=polar_angles, azimuthal_angles=azimuthal_angles,
                                                  angular_resolution=angular_resolution, precision=precision))
        ecs.append(extinction_cross_section(pw, parts, layer_system, precision=precision))
    return np.array(scs), np.array(ecs)


def extinction_cross_section(initial_field, particle_list, layer_system, precision='double'):
    """sigma_ext = -(P_bottom + P_top) / I_0 for a plane wave (both directions summed)."""
    i_top = layer_system.number_of_layers() - 1
    om = flds.angular_frequency(initial_field.vacuum_wavelength)
    k_bot = om * layer_system.refractive_indices[0]
    k_top = om * layer_system.refractive_indices[-1]
    pol, beta, alpha = initial_field.polarization, initial_field.polar_angle, initial_field.azimuthal_angle
    if np.cos(beta) > 0:
        iP, kP = 0, k_bot
    else:
        iP, kP = i_top, k_top
    nP = np.real(layer_system.refractive_indices[iP])
    kappa = np.sin(beta) * kP
    kvec = np.array([np.cos(alpha) * kappa, np.sin(alpha) * kappa, kP * np.cos(beta)])
    e = np.exp(1j * np.dot(kvec, np.array([0, 0, layer_system.reference_z(iP)])
                           - np.array(initial_field.reference_point)))
    A = initial_field.amplitude * e
    I0 = abs(A) ** 2 * nP / 2
    top, _ = scattered_field_pwe(initial_field.vacuum_wavelength, particle_list, layer_system, i_top, [kappa],
                                 [alpha], precision=precision)
    _, bot = scattered_field_pwe(initial_field.vacuum_wavelength, particle_list, layer_system, 0, [kappa], [alpha],
                                 precision=precision)
    _, ib = initial_field.plane_wave_expansion(layer_system, 0)
    it, _ = initial_field.plane_wave_expansion(layer_system, i_top)
    kzb = flds.k_z(k_parallel=kappa, k=k_bot)
    kzt = flds.k_z(k_parallel=kappa, k=k_top)
    Pb = 4 * np.pi ** 2 * kzb / om * np.real(np.squeeze(ib.coefficients[pol]) * np.conj(np.squeeze(bot.coefficients[pol])))
    Pt = 4 * np.pi ** 2 * kzt / om * np.real(np.squeeze(it.coefficients[pol]) * np.conj(np.squeeze(top.coefficients[pol])))
    return float(np.real(-Pb / I0) + np.real(-Pt / I0))
