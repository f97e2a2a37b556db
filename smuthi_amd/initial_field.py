"""Initial fields (initial_field.py:20-960) with the regular SWE coefficients
of all particles computed at once:

* PlaneWave: the discrete branch of pwe_to_swe_conversion
  (fields/transformations.py:77-130), vectorised on the host;
* GaussianBeam: its angular spectrum (a continuous plane wave expansion) is
  converted at every particle by the GPU kernel behind smh_pwe_to_swe;
* DipoleSource / DipoleCollection: the explicit coupling blocks from each
  dipole to every particle (smh_explicit_blocks) applied to the dipoles'
  outgoing l = 1 expansions.
"""
import ctypes
import warnings

import numpy as np

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import special as sf
from .expansions import PiecewiseFieldExpansion, PlaneWaveExpansion, SphericalWaveExpansion, layer_response


def _trapezoid_weights(x):
    """w with sum_j w_j f_j = numpy.trapz(f, x)."""
    x = np.asarray(x)
    d = np.diff(x)
    w = np.zeros(len(x), dtype=d.dtype if d.size else x.dtype)
    w[:-1] += d / 2
    w[1:] += d / 2
    return w


def pwe_to_swe(pwe, l_max, m_max, positions):
    """Regular SWE coefficients [n_points, blocksize(l_max, m_max)] of a plane
    wave expansion about each point (pwe_to_swe_conversion,
    transformations.py:77-130): a one-wave expansion on the host, a
    continuous one (trapezoid rule in alpha, then in kappa with the kappa
    Jacobian, times 4) by the FP64 GPU kernel."""
    positions = np.ascontiguousarray(np.asarray(positions, dtype=float).reshape(-1, 3))
    z = positions[:, 2]
    if np.any(z < pwe.lower_z) or np.any(z > pwe.upper_z):
        raise ValueError('reference point not inside domain of pwe validity')
    if len(pwe.k_parallel) == 1:
        return _discrete_pwe_to_swe(pwe, l_max, m_max, positions)
    kp = np.ascontiguousarray(pwe.k_parallel, dtype=np.complex128)
    kz = np.ascontiguousarray(pwe.k_z(), dtype=np.complex128)
    al = np.ascontiguousarray(pwe.azimuthal_angles, dtype=np.float64)
    # B-dagger [k][pol][n] at (kz / k, kappa / k)
    bd = np.ascontiguousarray(sf.transformation_arrays(l_max, m_max, kz / pwe.k, kp / pwe.k, True).transpose(2, 0, 1))
    w_k = np.ascontiguousarray(4 * kp * _trapezoid_weights(kp), dtype=np.complex128)
    w_a = np.ascontiguousarray(_trapezoid_weights(al), dtype=np.float64)
    g = np.ascontiguousarray(pwe.coefficients, dtype=np.complex128)
    pts = np.ascontiguousarray(positions - np.asarray(pwe.reference_point, dtype=float)[None, :])
    out = np.empty((len(pts), flds.blocksize(l_max, m_max)), dtype=np.complex128)
    ms = ctypes.c_double()
    ctx = cu.context()
    _lib.check(_lib.load().smh_pwe_to_swe(ctx, len(kp), _lib.dptr(kp), _lib.dptr(kz), _lib.dptr(w_k), len(al),
                                          _lib.dptr(al), _lib.dptr(w_a), _lib.dptr(g), _lib.dptr(bd), l_max, m_max,
                                          len(pts), _lib.dptr(pts), _lib.dptr(out), ctypes.byref(ms)), ctx,
               'smh_pwe_to_swe')
    pwe_to_swe.last_kernel_ms = ms.value
    return out


pwe_to_swe.last_kernel_ms = 0.0


def _regular_l1_wave_at_origin(tau, m):
    """Regular SVWF (tau, l = 1, m) at r = 0 as smuthi evaluates it there
    (vector_wave_functions.py:64-155 at theta = phi = 0, with j_1(x)/x -> 1/3
    and (x j_1(x))'/x -> 2/3): zero for tau = 0, else
    (2 j_1/x P_1^|m| e_r + (x j_1)'/x (tau_1^|m| e_theta + i m pi_1^|m| e_phi)) / 2
    with e_r = z, e_theta = x, e_phi = y."""
    if tau == 0:
        return np.zeros(3, complex)
    P, PI, TAU = sf.legendre_normalized(np.array(1.0), np.array(0.0), 1)
    am = abs(m)
    e_r, e_t, e_p = np.array([0.0, 0.0, 1.0]), np.array([1.0, 0.0, 0.0]), np.array([0.0, 1.0, 0.0])
    return (2 / 3 * complex(P[1][am]) * e_r + 2 / 3 * (complex(TAU[1][am]) * e_t + 1j * m * complex(PI[1][am]) * e_p)) / 2


def _system_vector(particle_list, layer_system, per_group):
    """Assemble a system vector from per_group(layer, l_max, m_max, positions) -> [n, blocksize]."""
    sizes = [flds.blocksize(p.l_max, p.m_max) for p in particle_list]
    offsets = np.concatenate([[0], np.cumsum(sizes)])
    out = np.zeros(offsets[-1], complex)
    groups = {}
    for j, p in enumerate(particle_list):
        key = (layer_system.layer_number(p.position[2]), p.l_max, p.m_max)
        groups.setdefault(key, []).append(j)
    for (i, l_max, m_max), idx in groups.items():
        pos = np.array([particle_list[j].position for j in idx], dtype=float)
        a = per_group(i, l_max, m_max, pos)
        rows = offsets[np.asarray(idx)][:, None] + np.arange(a.shape[1])[None, :]
        out[rows] = a
    return out


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
        def group(i, l_max, m_max, pos):
            up, dn = self.plane_wave_expansion(layer_system, i)
            return _discrete_pwe_to_swe(up, l_max, m_max, pos) + _discrete_pwe_to_swe(dn, l_max, m_max, pos)
        return _system_vector(particle_list, layer_system, group)

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


class GaussianBeam:
    """Gaussian beam (initial_field.py:157-269) with smuthi's constructor
    arguments. Its angular spectrum in the excitation layer is
    A w^2 / (4 pi) e^{-w^2 |k_par - k_G,par|^2 / 4} e^{-i k . (r_G - [0, 0, z_ref])}
    on the (kappa, alpha) grid, split into TE / TM by the angle to the beam's
    azimuth (the TM part negated for a beam from the top, as smuthi does)."""

    def __init__(self, vacuum_wavelength, polar_angle, azimuthal_angle, polarization, beam_waist,
                 k_parallel_array='default', azimuthal_angles_array='default', amplitude=1, reference_point=None):
        if polarization not in (0, 1):
            raise ValueError('polarization must be 0 (TE) or 1 (TM)')
        if np.isclose(np.cos(polar_angle), 0):
            raise ValueError('propagating waves not defined in the xy-plane')
        self.vacuum_wavelength = vacuum_wavelength
        self.polar_angle = polar_angle
        self.azimuthal_angle = azimuthal_angle
        self.polarization = polarization
        self.beam_waist = beam_waist
        self.k_parallel_array = k_parallel_array
        self.azimuthal_angles_array = azimuthal_angles_array
        self.amplitude = amplitude
        self.reference_point = list(reference_point) if reference_point else [0, 0, 0]

    def angular_frequency(self):
        return flds.angular_frequency(self.vacuum_wavelength)

    def get_k_parallel_array(self):
        if isinstance(self.k_parallel_array, str) and self.k_parallel_array == 'default':
            if flds.default_initial_field_k_parallel_array is None:
                raise ValueError('no default k_parallel array: pass k_parallel_array or run a Simulation first')
            return flds.default_initial_field_k_parallel_array
        return self.k_parallel_array

    def get_azimuthal_angles_array(self):
        if isinstance(self.azimuthal_angles_array, str) and self.azimuthal_angles_array == 'default':
            if flds.default_azimuthal_angles is None:
                flds.default_azimuthal_angles, flds.default_polar_angles = flds.angular_arrays()
            return flds.default_azimuthal_angles
        return self.azimuthal_angles_array

    def _source_side(self, layer_system):
        """(layer the beam enters from, +1 travelling up / -1 down): the bottom
        half-space for cos(polar angle) > 0, else the top one."""
        if np.cos(self.polar_angle) > 0:
            return 0, 1
        return layer_system.number_of_layers() - 1, -1

    def _spectrum(self, k, kappa, alpha, kz, z_ref):
        """Gaussian angular spectrum on the (kappa, alpha) grid of the source
        layer with wavenumber k and anchor height z_ref: the waist-w Gaussian
        around the beam's in-plane wavevector times the shift to the beam's
        reference point."""
        centre = k * np.sin(self.polar_angle) * np.array([np.cos(self.azimuthal_angle),
                                                          np.sin(self.azimuthal_angle)])
        qx, qy = kappa * np.cos(alpha), kappa * np.sin(alpha)
        w2 = self.beam_waist ** 2
        shift = np.asarray(self.reference_point, dtype=float) - np.array([0.0, 0.0, z_ref])
        envelope = np.exp(-0.25 * w2 * ((qx - centre[0]) ** 2 + (qy - centre[1]) ** 2))
        return (self.amplitude * w2 / (4 * np.pi)) * envelope * np.exp(-1j * (qx * shift[0] + qy * shift[1]
                                                                               + kz * shift[2]))

    def plane_wave_expansion(self, layer_system, i, k_parallel_array=None, azimuthal_angles_array=None):
        """(upgoing, downgoing) PWEs of the beam in layer i (initial_field.py:167-233): the beam's
        spectrum in its source half-space, polarised TE / TM by its angle to the beam's azimuth (TM
        negated for a beam travelling down), carried to layer i by the layer response."""
        kappa = np.asarray(self.get_k_parallel_array() if k_parallel_array is None else k_parallel_array)
        alpha = np.asarray(self.get_azimuthal_angles_array() if azimuthal_angles_array is None
                           else azimuthal_angles_array)
        src, direction = self._source_side(layer_system)
        n_src = layer_system.refractive_indices[src]
        if np.imag(n_src):
            warnings.warn('beam coming from absorbing medium')
        z_src = layer_system.reference_z(src)
        beam = PlaneWaveExpansion(n_src * self.angular_frequency(), kappa, alpha,
                                  'upgoing' if direction > 0 else 'downgoing', [0, 0, z_src],
                                  layer_system.lower_zlimit(src), layer_system.upper_zlimit(src))
        kap, alp = np.meshgrid(beam.k_parallel, beam.azimuthal_angles, indexing='ij')
        g = self._spectrum(beam.k, kap, alp, beam.k_z()[:, None], z_src)
        rel = alp - self.azimuthal_angle + 0.5 * np.pi * self.polarization
        beam.coefficients[0] = g * np.cos(rel)
        beam.coefficients[1] = direction * g * np.sin(rel)
        up, dn = layer_response(layer_system, beam, src, i)
        if src == i:
            if direction > 0:
                up = up + beam
            else:
                dn = dn + beam
        return up, dn

    def spherical_wave_expansion(self, particle, layer_system):
        """Regular SWE of the beam at one particle (initial_field.py:92-106)."""
        i = layer_system.layer_number(particle.position[2])
        swe = SphericalWaveExpansion(flds.angular_frequency(self.vacuum_wavelength)
                                     * layer_system.refractive_indices[i], particle.l_max, particle.m_max,
                                     'regular', particle.position, layer_system.lower_zlimit(i),
                                     layer_system.upper_zlimit(i))
        swe.coefficients = self.swe_coefficients([particle], layer_system)
        return swe

    def swe_coefficients(self, particle_list, layer_system):
        """Regular SWE coefficients of all particles as one system vector: per
        layer the two PWEs are built once and converted at every particle of
        that layer on the GPU."""
        cache = {}

        def group(i, l_max, m_max, pos):
            if i not in cache:
                cache[i] = self.plane_wave_expansion(layer_system, i)
            up, dn = cache[i]
            return pwe_to_swe(up, l_max, m_max, pos) + pwe_to_swe(dn, l_max, m_max, pos)
        return _system_vector(particle_list, layer_system, group)

    def piecewise_field_expansion(self, layer_system):
        pfe = PiecewiseFieldExpansion()
        for i in range(layer_system.number_of_layers()):
            up, dn = self.plane_wave_expansion(layer_system, i)
            pfe.expansion_list += [up, dn]
        return pfe

    def electric_field(self, x, y, z, layer_system, precision='double'):
        return self.piecewise_field_expansion(layer_system).electric_field(x, y, z, precision=precision)

    def magnetic_field(self, x, y, z, layer_system, precision='double'):
        return self.piecewise_field_expansion(layer_system).magnetic_field(x, y, z, self.vacuum_wavelength,
                                                                          precision=precision)

    def initial_intensity(self, layer_system):
        """Far-field intensity of the incoming beam (initial_field.py:253-269)."""
        from .far_field import pwe_to_ff_conversion
        if np.cos(self.polar_angle) > 0:
            return pwe_to_ff_conversion(self.vacuum_wavelength, self.plane_wave_expansion(layer_system, 0)[0])
        top = layer_system.number_of_layers() - 1
        return pwe_to_ff_conversion(self.vacuum_wavelength, self.plane_wave_expansion(layer_system, top)[1])


class DipoleSource:
    """Point dipole source (initial_field.py:331-640) with smuthi's constructor
    arguments; current density j = -i omega mu at the dipole position."""

    def __init__(self, vacuum_wavelength, dipole_moment, position, k_parallel_array='default',
                 azimuthal_angles_array='default'):
        self.vacuum_wavelength = vacuum_wavelength
        self.dipole_moment = list(dipole_moment)
        self.position = list(position)
        self.k_parallel_array = k_parallel_array
        self.azimuthal_angles_array = azimuthal_angles_array

    def angular_frequency(self):
        return flds.angular_frequency(self.vacuum_wavelength)

    def get_k_parallel_array(self):
        if isinstance(self.k_parallel_array, str) and self.k_parallel_array == 'default':
            if flds.default_initial_field_k_parallel_array is None:
                raise ValueError('no default k_parallel array: pass k_parallel_array or run a Simulation first')
            return flds.default_initial_field_k_parallel_array
        return self.k_parallel_array

    def current(self):
        return [-1j * self.angular_frequency() * d for d in self.dipole_moment]

    def outgoing_spherical_wave_expansion(self, layer_system):
        """The dipole field as an outgoing l_max = m_max = 1 SWE (initial_field.py:366-388):
        b_(tau, 1, m) = i k / pi * i omega * N^(1)_(tau, 1, -m)(0) . j."""
        i = layer_system.layer_number(self.position[2])
        k = layer_system.refractive_indices[i] * self.angular_frequency()
        swe = SphericalWaveExpansion(k, 1, 1, 'outgoing', self.position, layer_system.lower_zlimit(i),
                                     layer_system.upper_zlimit(i))
        j = np.asarray(self.current())
        om = self.angular_frequency()
        b = np.zeros(flds.blocksize(1, 1), complex)
        for tau in range(2):
            for m in (-1, 0, 1):
                e = _regular_l1_wave_at_origin(tau, -m)
                b[flds.multi_to_single_index(tau, 1, m, 1, 1)] = 1j * k / np.pi * 1j * om * (e @ j)
        swe.coefficients = b
        return swe

    def swe_coefficients(self, particle_list, layer_system):
        return DipoleCollection.of([self], self.vacuum_wavelength).swe_coefficients(particle_list, layer_system)

    def spherical_wave_expansion(self, particle, layer_system):
        i = layer_system.layer_number(particle.position[2])
        swe = SphericalWaveExpansion(self.angular_frequency() * layer_system.refractive_indices[i], particle.l_max,
                                     particle.m_max, 'regular', particle.position)
        swe.coefficients = self.swe_coefficients([particle], layer_system)
        return swe


class DipoleCollection:
    """Set of dipole sources (initial_field.py:642-960); the particles' regular
    SWE is the sum of every dipole's (the per-dipole branch,
    initial_field.py:684-709), formed from one GPU build of the blocks
    between all particles and all dipoles."""

    def __init__(self, vacuum_wavelength, k_parallel_array='default', azimuthal_angles_array='default',
                 angular_resolution=None, compute_swe_by_pwe=False, compute_dissipated_power_by_pwe=False):
        self.vacuum_wavelength = vacuum_wavelength
        self.k_parallel_array = k_parallel_array
        self.azimuthal_angles_array = azimuthal_angles_array
        self.compute_swe_by_pwe = compute_swe_by_pwe
        self.dipole_list = []

    @classmethod
    def of(cls, dipoles, vacuum_wavelength):
        c = cls(vacuum_wavelength, k_parallel_array=dipoles[0].k_parallel_array)
        c.dipole_list = list(dipoles)
        return c

    def append(self, dipole):
        if dipole.vacuum_wavelength != self.vacuum_wavelength:
            raise ValueError('dipole and collection differ in wavelength')
        self.dipole_list.append(dipole)

    def angular_frequency(self):
        return flds.angular_frequency(self.vacuum_wavelength)

    def get_k_parallel_array(self):
        if isinstance(self.k_parallel_array, str) and self.k_parallel_array == 'default':
            if flds.default_initial_field_k_parallel_array is None:
                raise ValueError('no default k_parallel array: pass k_parallel_array or run a Simulation first')
            return flds.default_initial_field_k_parallel_array
        return self.k_parallel_array

    def swe_coefficients(self, particle_list, layer_system):
        """Regular SWE coefficients of all particles as one system vector:
        a_i = sum_d (W_direct + W_layer)(particle i <- dipole d) b_d."""
        from .linear_system import explicit_blocks
        if not self.dipole_list:
            raise ValueError('no dipoles')
        l_max = max(p.l_max for p in particle_list)
        m_max = max(max(p.m_max for p in particle_list), 1)
        B = flds.blocksize(l_max, m_max)
        recv = [p.position for p in particle_list]
        emit = [d.position for d in self.dipole_list]
        # each dipole's outgoing l = 1 SWE in the padded (l_max, m_max) order
        bp = np.zeros((len(self.dipole_list), B), complex)
        for q, d in enumerate(self.dipole_list):
            b = d.outgoing_spherical_wave_expansion(layer_system).coefficients
            for tau in range(2):
                for m in (-1, 0, 1):
                    bp[q, flds.multi_to_single_index(tau, 1, m, l_max, m_max)] = b[
                        flds.multi_to_single_index(tau, 1, m, 1, 1)]
        a_pad = explicit_blocks(self.vacuum_wavelength, recv, emit, layer_system, self.get_k_parallel_array(), l_max,
                                m_max, x=bp)
        out = []
        for i, p in enumerate(particle_list):
            idx = [flds.multi_to_single_index(t, l, m, l_max, m_max)
                   for t in range(2) for l in range(1, p.l_max + 1) for m in range(-min(l, p.m_max), min(l, p.m_max) + 1)]
            out.append(a_pad[i, idx])
        return np.concatenate(out)

    def spherical_wave_expansion(self, particle, layer_system):
        i = layer_system.layer_number(particle.position[2])
        swe = SphericalWaveExpansion(self.angular_frequency() * layer_system.refractive_indices[i], particle.l_max,
                                     particle.m_max, 'regular', particle.position)
        swe.coefficients = self.swe_coefficients([particle], layer_system)
        return swe
