"""Plane wave expansion of the scattered field with the particle sum on the
GPU (postprocessing/scattered_field.py:13-150): for every layer of the
scatterers, the outgoing SWE -> PWE conversion
(fields/transformations.py:133-226) of all particles is summed in one
libsmh launch (smh_pwe_assemble); the direct part and the layer response
(layers.py:141-191) are combined on the host.
"""
import ctypes

import numpy as np

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import special as sf
from .expansions import PiecewiseFieldExpansion, PlaneWaveExpansion, SphericalWaveSum, layer_response


def _coefficients(particle):
    swe = particle.scattered_field
    if swe is None:
        raise ValueError('particle has no scattered field (solve the linear system first)')
    return np.asarray(getattr(swe, 'coefficients', swe), dtype=complex)


def _assemble(vacuum_wavelength, particles, layer_system, i_swe, k_parallel, azimuthal_angles, precision):
    """(g_up, g_dn), each [2, n_k, n_a], of the particles of layer i_swe at reference [0, 0, z_ref(i_swe)]."""
    om = flds.angular_frequency(vacuum_wavelength)
    k = complex(om * layer_system.refractive_indices[i_swe])
    l_max = max(p.l_max for p in particles)
    m_max = max(p.m_max for p in particles)
    B = flds.blocksize(l_max, m_max)
    coef = np.zeros((len(particles), B), complex)
    # particles grouped by their own (l_max, m_max): one index map and one array copy per group
    groups = {}
    for i, p in enumerate(particles):
        groups.setdefault((p.l_max, p.m_max), []).append(i)
    for (lp, mp), idx in groups.items():
        dst = [flds.multi_to_single_index(tau, l, m, l_max, m_max)
               for tau in range(2) for l in range(1, lp + 1) for m in range(-min(l, mp), min(l, mp) + 1)]
        src = [flds.multi_to_single_index(tau, l, m, lp, mp)
               for tau in range(2) for l in range(1, lp + 1) for m in range(-min(l, mp), min(l, mp) + 1)]
        block = np.stack([_coefficients(particles[i]) for i in idx])
        coef[np.asarray(idx)[:, None], np.asarray(dst)[None, :]] = block[:, src]
    kp = np.ascontiguousarray(np.array(k_parallel, ndmin=1), dtype=complex)
    kz = flds.k_z(k_parallel=kp, k=k).astype(complex)
    # B_{n,pol}(+-kz/k, kappa/k) / (2 pi kz k): btab[k][dir][pol][n]
    bu = sf.transformation_arrays(l_max, m_max, kz / k, kp / k, False)
    bd = sf.transformation_arrays(l_max, m_max, -kz / k, kp / k, False)
    den = 2 * np.pi * kz * k
    btab = np.ascontiguousarray(np.stack([bu, bd]).transpose(3, 0, 1, 2) / den[:, None, None, None])
    al = np.ascontiguousarray(np.array(azimuthal_angles, ndmin=1), dtype=np.float64)
    xyz = np.ascontiguousarray([p.position for p in particles], dtype=np.float64)
    g = np.zeros((2, 2, len(kp), len(al)), complex)
    ms = ctypes.c_double()
    prec = _lib.FP64 if precision in ('double', 'double precision') else _lib.FP32
    _lib.check(_lib.load().smh_pwe_assemble(cu.context(), len(particles), _lib.dptr(xyz), l_max, m_max,
                                             _lib.dptr(np.ascontiguousarray(coef)),
                                             float(layer_system.reference_z(i_swe)), len(kp), _lib.dptr(kp),
                                             _lib.dptr(np.ascontiguousarray(kz)), _lib.dptr(btab), len(al),
                                             _lib.dptr(al), _lib.dptr(g), prec, ctypes.byref(ms)),
               cu.context(), 'smh_pwe_assemble')
    return g[0], g[1], k, kp, al, ms.value


def scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, layer_number, k_parallel='default',
                        azimuthal_angles='default', include_direct=True, include_layer_response=True,
                        precision='single'):
    """(upgoing, downgoing) PlaneWaveExpansion of the scattered field in layer_number
    (postprocessing/scattered_field.py:80-150)."""
    if isinstance(k_parallel, str) and k_parallel == 'default':
        k_parallel = flds.default_Sommerfeld_k_parallel_array
    if isinstance(azimuthal_angles, str) and azimuthal_angles == 'default':
        azimuthal_angles = flds.default_azimuthal_angles
    om = flds.angular_frequency(vacuum_wavelength)
    k = om * layer_system.refractive_indices[layer_number]
    z = layer_system.reference_z(layer_number)
    vb = (layer_system.lower_zlimit(layer_number), layer_system.upper_zlimit(layer_number))
    up = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'upgoing', [0, 0, z], vb[0], vb[1])
    dn = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'downgoing', [0, 0, z], vb[0], vb[1])
    groups = {}
    for p in particle_list:
        groups.setdefault(layer_system.layer_number(p.position[2]), []).append(p)
    scattered_field_pwe.last_kernel_ms = 0.0
    for i_swe, parts in sorted(groups.items()):
        direct = i_swe == layer_number and include_direct
        if not (direct or include_layer_response):
            continue
        gu, gd, k_swe, kp, al, ms = _assemble(vacuum_wavelength, parts, layer_system, i_swe, k_parallel,
                                              azimuthal_angles, precision)
        scattered_field_pwe.last_kernel_ms += ms
        ref = [0, 0, layer_system.reference_z(i_swe)]
        zs = [p.position[2] for p in parts]
        pu = PlaneWaveExpansion(k_swe, kp, al, 'upgoing', ref, max(zs), layer_system.upper_zlimit(i_swe))
        pd = PlaneWaveExpansion(k_swe, kp, al, 'downgoing', ref, layer_system.lower_zlimit(i_swe), min(zs))
        pu.coefficients, pd.coefficients = gu, gd
        if direct:
            up, dn = up + pu, dn + pd
        if include_layer_response:
            ru, rd = layer_response(layer_system, (pu, pd), i_swe, layer_number)
            up, dn = up + ru, dn + rd
    return up, dn


def scattered_field_piecewise_expansion(vacuum_wavelength, particle_list, layer_system, k_parallel='default',
                                        azimuthal_angles='default', layer_numbers=None, precision='single'):
    """Scattered field as a PiecewiseFieldExpansion
    (postprocessing/scattered_field.py:13-77): per layer the layer-mediated
    upgoing (not in the bottom layer) and downgoing (not in the top layer)
    plane wave expansions, plus the direct field, the particles' outgoing
    spherical wave expansions, summed on the GPU per layer
    (expansions.SphericalWaveSum); points inside a particle get no field."""
    if layer_numbers is None:
        layer_numbers = range(layer_system.number_of_layers())
    nl = layer_system.number_of_layers()
    sfld = PiecewiseFieldExpansion()
    for i in layer_numbers:
        up, dn = scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, i, k_parallel, azimuthal_angles,
                                     include_direct=False, include_layer_response=True, precision=precision)
        if i > 0:
            sfld.expansion_list.append(up)
        if i < nl - 1:
            sfld.expansion_list.append(dn)
    by_layer = {}
    for p in particle_list:
        by_layer.setdefault(layer_system.layer_number(p.position[2]), []).append(p)
    for i, parts in sorted(by_layer.items()):
        k = flds.angular_frequency(vacuum_wavelength) * layer_system.refractive_indices[i]
        sfld.expansion_list.append(SphericalWaveSum(k, parts, layer_system.lower_zlimit(i),
                                                    layer_system.upper_zlimit(i)))
    for p in particle_list:
        sfld.validity_conditions.append(p.is_outside)
    return sfld
