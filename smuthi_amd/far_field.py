"""Extinction cross section (postprocessing/far_field.py:517-654) from the
scattered-field plane wave expansion at the incident direction, assembled
on the GPU."""
import numpy as np

from . import fields as flds
from .scattered_field import scattered_field_pwe


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
