"""Oracle system level: spheres, plane-wave initial field, plane wave
expansions and their fields, SWE<->PWE conversion, the linear system
(1 - T W) b = T a, GMRES / LU solves and the extinction cross section.

Pinning (tests/test_oracle_system.py):
  * full single-sphere solve (PlaneWave RHS, Mie T, explicit layer-mediated
    self coupling, LU): prototype coefficients b0, b10, b21 of
    tests/system_tests/test_single_sphere_on_substrate.py:56-62 (rel 1e-4).
  * GMRES vs LU: tests/unit_tests/test_gmres.py:57 (rtol 1e-5), plus the true
    residual ||b - M x|| <= tol ||b||.
  * one sphere in a homogeneous medium: b = Q a (Mie closed form) and
    sigma_ext equals the Mie extinction formula.
  * PWE fields: a discrete (single-kappa) PWE is g e^{ik.r} exactly (closed
    form); the double-trapezoid sum equals a brute-force quadrature written as
    an explicit loop on tiny grids; E and H of a plane wave satisfy
    H = k x E / omega.
  * swe_to_pwe / pwe_to_swe: regular SWE of a plane wave re-expanded as a PWE
    reproduces the plane wave (round trip), as in
    test_field_expansion_transformation.py:44-61.
"""

import numpy as np
import scipy.linalg
import scipy.sparse.linalg

from . import basics as fb
from . import layers as lay
from . import coupling as cp


class Sphere:
    """Sphere particle (particles.py:97-147)."""

    def __init__(self, position, refractive_index, radius, l_max, m_max=None):
        self.position = [float(c) for c in position]
        self.refractive_index = refractive_index
        self.radius = radius
        self.l_max = l_max
        self.m_max = l_max if m_max is None else m_max
        self.initial_field = None
        self.scattered_field = None

    def is_outside(self, x, y, z):
        return ((x - self.position[0]) ** 2 + (y - self.position[1]) ** 2
                + (z - self.position[2]) ** 2 > self.radius ** 2)

    def t_matrix(self, vacuum_wavelength, n_medium):
        k_m = 2 * np.pi / vacuum_wavelength * n_medium
        k_p = 2 * np.pi / vacuum_wavelength * self.refractive_index
        return lay.t_matrix_sphere(k_m, k_p, self.radius, self.l_max, self.m_max)


class PlaneWaveExpansion:
    """PWE g_j(kappa, alpha) (fields/expansions.py:466-633)."""

    def __init__(self, k, k_parallel, azimuthal_angles, kind, reference_point, lower_z=-np.inf, upper_z=np.inf):
        self.k = k
        self.k_parallel = np.array(k_parallel, ndmin=1)
        self.azimuthal_angles = np.array(azimuthal_angles, ndmin=1)
        self.kind = kind
        self.reference_point = list(reference_point)
        self.lower_z = lower_z
        self.upper_z = upper_z
        self.coefficients = np.zeros((2, len(self.k_parallel), len(self.azimuthal_angles)), dtype=complex)

    def k_z(self):
        kz = fb.k_z(k_parallel=self.k_parallel, k=self.k)
        return kz if self.kind == 'upgoing' else -kz

    def grids(self):
        kp, al = np.meshgrid(self.k_parallel, self.azimuthal_angles, indexing='ij')
        kz = fb.k_z(k_parallel=kp, k=self.k)
        return kp, al, (kz if self.kind == 'upgoing' else -kz)

    def valid(self, x, y, z):
        if self.upper_z == float('inf'):
            return np.asarray(z) >= self.lower_z
        return np.logical_and(np.asarray(z) >= self.lower_z, np.asarray(z) < self.upper_z)

    def __add__(self, other):
        s = PlaneWaveExpansion(self.k, self.k_parallel, self.azimuthal_angles, self.kind, self.reference_point,
                               max(self.lower_z, other.lower_z), min(self.upper_z, other.upper_z))
        s.coefficients = self.coefficients + other.coefficients
        return s

    def _integrands(self, field):
        kp, al, kz = self.grids()
        g0, g1 = self.coefficients[0], self.coefficients[1]
        if field == 'E':
            fx = -np.sin(al) * g0 + np.cos(al) * kz / self.k * g1
            fy = np.cos(al) * g0 + np.sin(al) * kz / self.k * g1
            fz = -kp / self.k * g1
        else:
            fx = -kz * np.cos(al) * g0 - np.sin(al) * self.k * g1
            fy = -kz * np.sin(al) * g0 + np.cos(al) * self.k * g1
            fz = kp * g0
        return kp, al, kz, (fx, fy, fz)

    def field(self, x, y, z, field='E', omega=1.0):
        """Double trapezoid (alpha inner, kappa outer) in FP64
        (fields/expansions.py:635-948 CPU path with 'double precision')."""
        x, y, z = (np.atleast_1d(np.asarray(a, dtype=float)) for a in (x, y, z))
        out = [np.zeros(x.shape, dtype=complex) for _ in range(3)]
        vld = self.valid(x, y, z)
        xr = x[vld] - self.reference_point[0]
        yr = y[vld] - self.reference_point[1]
        zr = z[vld] - self.reference_point[2]
        kp, al, kz, f = self._integrands(field)
        kx, ky = kp * np.cos(al), kp * np.sin(al)
        res = [np.zeros(xr.shape, dtype=complex) for _ in range(3)]
        for i in range(xr.size):
            e = np.exp(1j * (kx * xr[i] + ky * yr[i] + kz * zr[i]))
            for c in range(3):
                if len(self.k_parallel) > 1:
                    inner = np.trapezoid(f[c] * e, self.azimuthal_angles, axis=1)
                    res[c][i] = np.trapezoid(inner * self.k_parallel, self.k_parallel) / omega
                else:
                    res[c][i] = (f[c] * e)[0, 0] / omega
        for c in range(3):
            out[c][vld] = res[c]
        return tuple(out)

    def electric_field(self, x, y, z):
        return self.field(x, y, z, 'E', 1.0)

    def magnetic_field(self, x, y, z, vacuum_wavelength):
        return self.field(x, y, z, 'H', fb.angular_frequency(vacuum_wavelength))


def layer_response(layer_system, pwe, from_layer, to_layer):
    """LayerSystem.response (layers.py:141-191)."""
    if isinstance(pwe, tuple):
        u0, d0 = layer_response(layer_system, pwe[0], from_layer, to_layer)
        u1, d1 = layer_response(layer_system, pwe[1], from_layer, to_layer)
        return u0 + u1, d0 + d1
    omega = pwe.k / layer_system.refractive_indices[from_layer]
    k_to = omega * layer_system.refractive_indices[to_layer]
    ref = [0, 0, layer_system.reference_z(to_layer)]
    loz, upz = layer_system.lower_zlimit(to_layer), layer_system.upper_zlimit(to_layer)
    up = PlaneWaveExpansion(k_to, pwe.k_parallel, pwe.azimuthal_angles, 'upgoing', ref, loz, upz)
    dn = PlaneWaveExpansion(k_to, pwe.k_parallel, pwe.azimuthal_angles, 'downgoing', ref, loz, upz)
    for pol in range(2):
        L = lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                            pwe.k_parallel, omega, from_layer, to_layer)
        col = 0 if pwe.kind == 'upgoing' else 1
        up.coefficients[pol] = L[0, col, :][:, None] * pwe.coefficients[pol]
        dn.coefficients[pol] = L[1, col, :][:, None] * pwe.coefficients[pol]
    return up, dn


class PlaneWave:
    """Plane-wave initial field (initial_field.py:284-328)."""

    def __init__(self, vacuum_wavelength, polar_angle, azimuthal_angle, polarization, amplitude=1,
                 reference_point=None):
        self.vacuum_wavelength = vacuum_wavelength
        self.polar_angle = polar_angle
        self.azimuthal_angle = azimuthal_angle
        self.polarization = polarization
        self.amplitude = amplitude
        self.reference_point = list(reference_point) if reference_point else [0, 0, 0]

    def plane_wave_expansion(self, layer_system, i):
        if np.cos(self.polar_angle) > 0:
            iP, kind = 0, 'upgoing'
        else:
            iP, kind = layer_system.number_of_layers() - 1, 'downgoing'
        niP = layer_system.refractive_indices[iP]
        neff = np.sin([self.polar_angle]) * niP
        alpha = np.array([self.azimuthal_angle])
        om = fb.angular_frequency(self.vacuum_wavelength)
        k_iP = niP * om
        kPx = k_iP * np.sin(self.polar_angle) * np.cos(self.azimuthal_angle)
        kPy = k_iP * np.sin(self.polar_angle) * np.sin(self.azimuthal_angle)
        kPz = k_iP * np.cos(self.polar_angle)
        z_iP = layer_system.reference_z(iP)
        amp = self.amplitude * np.exp(-1j * (kPx * self.reference_point[0] + kPy * self.reference_point[1]
                                             + kPz * (self.reference_point[2] - z_iP)))
        pwe = PlaneWaveExpansion(k_iP, neff * om, alpha, kind, [0, 0, z_iP],
                                 layer_system.lower_zlimit(iP), layer_system.upper_zlimit(iP))
        pwe.coefficients[self.polarization, 0, 0] = amp
        up, dn = layer_response(layer_system, pwe, iP, i)
        if iP == i:
            if kind == 'upgoing':
                up = up + pwe
            else:
                dn = dn + pwe
        return up, dn


def pwe_to_swe_conversion(pwe, l_max, m_max, reference_point):
    """Regular SWE coefficients of a PWE (fields/transformations.py:77-130)."""
    coeffs = np.zeros(fb.blocksize(l_max, m_max), dtype=complex)
    kp, al, kz = pwe.grids()
    kx, ky = kp * np.cos(al), kp * np.sin(al)
    d = np.array(reference_point) - np.array(pwe.reference_point)
    ejk = np.exp(1j * (kx * d[0] + ky * d[1] + kz * d[2]))
    gej = pwe.coefficients * ejk[None]
    kzv = pwe.k_z()
    _, pilm, taulm = fb.legendre_normalized(kzv / pwe.k, pwe.k_parallel / pwe.k, l_max)
    for m in range(-m_max, m_max + 1):
        emj = np.exp(-1j * m * pwe.azimuthal_angles)[None, None, :] * gej
        for l in range(max(1, abs(m)), l_max + 1):
            for tau in range(2):
                integ = np.zeros(kp.shape, dtype=complex)
                for pol in range(2):
                    Bd = fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pilm, taulm_list=taulm,
                                                            dagger=True)
                    integ += Bd[:, None] * emj[pol]
                if len(pwe.k_parallel) > 1:
                    an = np.trapezoid(np.trapezoid(integ, pwe.azimuthal_angles) * pwe.k_parallel, pwe.k_parallel) * 4
                else:
                    an = integ * 4
                coeffs[fb.multi_to_single_index(tau, l, m, l_max, m_max)] = np.squeeze(an)
    return coeffs


def initial_field_coefficients(initial_field, particle, layer_system):
    """Regular SWE of the initial field at a particle (initial_field.py:92-106)."""
    i = layer_system.layer_number(particle.position[2])
    up, dn = initial_field.plane_wave_expansion(layer_system, i)
    return (pwe_to_swe_conversion(up, particle.l_max, particle.m_max, particle.position)
            + pwe_to_swe_conversion(dn, particle.l_max, particle.m_max, particle.position))


def swe_to_pwe_conversion(coeffs, l_max, m_max, k, position, k_parallel, azimuthal_angles, layer_system,
                          layer_number=None, layer_system_mediated=False):
    """Outgoing SWE -> (upgoing, downgoing) PWE (fields/transformations.py:133-226)."""
    k_parallel = np.array(k_parallel, ndmin=1)
    i_swe = layer_system.layer_number(position[2])
    if layer_number is None and not layer_system_mediated:
        layer_number = i_swe
    ref = [0, 0, layer_system.reference_z(i_swe)]
    up = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'upgoing', ref, position[2],
                            layer_system.upper_zlimit(layer_number))
    dn = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'downgoing', ref,
                            layer_system.lower_zlimit(layer_number), position[2])
    kp, al, kz_up = up.grids()
    _, _, kz_dn = dn.grids()
    kx, ky = kp * np.cos(al), kp * np.sin(al)
    d = np.array(ref) - np.array(position)
    e_up = np.exp(1j * (kx * d[0] + ky * d[1] + kz_up * d[2]))
    e_dn = np.exp(1j * (kx * d[0] + ky * d[1] + kz_dn * d[2]))
    kzv = up.k_z()
    _, pi_up, tau_up = fb.legendre_normalized(up.k_z() / k, k_parallel / k, l_max)
    _, pi_dn, tau_dn = fb.legendre_normalized(dn.k_z() / k, k_parallel / k, l_max)
    for m in range(-m_max, m_max + 1):
        eima = np.exp(1j * m * up.azimuthal_angles)
        for pol in range(2):
            bu = np.zeros(len(k_parallel), dtype=complex)
            bd = np.zeros(len(k_parallel), dtype=complex)
            for l in range(max(1, abs(m)), l_max + 1):
                for tau in range(2):
                    c = coeffs[fb.multi_to_single_index(tau, l, m, l_max, m_max)]
                    bu += c * fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_up, taulm_list=tau_up)
                    bd += c * fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_dn, taulm_list=tau_dn)
            up.coefficients[pol] += bu[:, None] * eima[None, :]
            dn.coefficients[pol] += bd[:, None] * eima[None, :]
    up.coefficients = up.coefficients / (2 * np.pi * kzv[None, :, None] * k) * e_up[None]
    dn.coefficients = dn.coefficients / (2 * np.pi * kzv[None, :, None] * k) * e_dn[None]
    if layer_system_mediated:
        up, dn = layer_response(layer_system, (up, dn), i_swe, layer_number)
    return up, dn


def scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, layer_number, k_parallel, azimuthal_angles,
                        include_direct=True, include_layer_response=True):
    """PWE of the scattered field in one layer (postprocessing/scattered_field.py:80-150)."""
    om = fb.angular_frequency(vacuum_wavelength)
    k = om * layer_system.refractive_indices[layer_number]
    z = layer_system.reference_z(layer_number)
    vb = (layer_system.lower_zlimit(layer_number), layer_system.upper_zlimit(layer_number))
    up = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'upgoing', [0, 0, z], vb[0], vb[1])
    dn = PlaneWaveExpansion(k, k_parallel, azimuthal_angles, 'downgoing', [0, 0, z], vb[0], vb[1])
    for p in particle_list:
        ip = layer_system.layer_number(p.position[2])
        kp_ = om * layer_system.refractive_indices[ip]
        if ip == layer_number and include_direct:
            pu, pd = swe_to_pwe_conversion(p.scattered_field, p.l_max, p.m_max, kp_, p.position, k_parallel,
                                           azimuthal_angles, layer_system)
            up, dn = up + pu, dn + pd
        if include_layer_response:
            pu, pd = swe_to_pwe_conversion(p.scattered_field, p.l_max, p.m_max, kp_, p.position, k_parallel,
                                           azimuthal_angles, layer_system, layer_number, True)
            up, dn = up + pu, dn + pd
    return up, dn


def extinction_cross_section(initial_field, particle_list, layer_system):
    """Extinction cross section, both directions (postprocessing/far_field.py:517-654)."""
    i_top = layer_system.number_of_layers() - 1
    om = fb.angular_frequency(initial_field.vacuum_wavelength)
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
    top, _ = scattered_field_pwe(initial_field.vacuum_wavelength, particle_list, layer_system, i_top, kappa, alpha)
    _, bot = scattered_field_pwe(initial_field.vacuum_wavelength, particle_list, layer_system, 0, kappa, alpha)
    _, ib = initial_field.plane_wave_expansion(layer_system, 0)
    it, _ = initial_field.plane_wave_expansion(layer_system, i_top)
    kzb = fb.k_z(k_parallel=kappa, k=k_bot)
    kzt = fb.k_z(k_parallel=kappa, k=k_top)
    Pb = 4 * np.pi ** 2 * kzb / om * (np.squeeze(ib.coefficients[pol]) * np.conj(np.squeeze(bot.coefficients[pol]))).real
    Pt = 4 * np.pi ** 2 * kzt / om * (np.squeeze(it.coefficients[pol]) * np.conj(np.squeeze(top.coefficients[pol]))).real
    return float(np.real(-Pb / I0) + np.real(-Pt / I0))


class LinearSystemOracle:
    """(1 - T W) b = T a with the explicit or lookup coupling
    (linearsystem/linear_system.py:42-344, 388-438)."""

    def __init__(self, particle_list, initial_field, layer_system, k_parallel):
        self.particles = particle_list
        self.initial_field = initial_field
        self.layer_system = layer_system
        self.k_parallel = k_parallel
        self.part, self.nlk, self.m, self.offs = cp.system_index_maps(particle_list)
        self.n_unk = int(self.offs[-1])
        wl = initial_field.vacuum_wavelength
        self.t_blocks = []
        self.a = np.zeros(self.n_unk, dtype=complex)
        for i, p in enumerate(particle_list):
            iS = layer_system.layer_number(p.position[2])
            self.t_blocks.append(p.t_matrix(wl, layer_system.refractive_indices[iS]))
            self.a[self.offs[i]:self.offs[i + 1]] = initial_field_coefficients(initial_field, p, layer_system)

    def apply_t(self, v):
        out = np.zeros(v.shape, dtype=complex)
        for i, t in enumerate(self.t_blocks):
            s = slice(self.offs[i], self.offs[i + 1])
            out[s] = t @ v[s]
        return out

    def right_hand_side(self):
        return self.apply_t(self.a)

    def explicit_coupling(self):
        wl = self.initial_field.vacuum_wavelength
        W = np.zeros((self.n_unk, self.n_unk), dtype=complex)
        for i1, p1 in enumerate(self.particles):
            for i2, p2 in enumerate(self.particles):
                blk = cp.direct_coupling_block(wl, p1, p2, self.layer_system)
                if not self.layer_system.is_degenerate():
                    blk = blk + cp.layer_mediated_coupling_block(wl, p1, p2, self.layer_system, self.k_parallel)
                W[self.offs[i1]:self.offs[i1 + 1], self.offs[i2]:self.offs[i2 + 1]] = blk
        return W

    def master_dense(self, W):
        T = scipy.linalg.block_diag(*self.t_blocks)
        return np.eye(self.n_unk) - T @ W

    def solve_lu(self, W):
        return np.linalg.solve(self.master_dense(W), self.right_hand_side())

    def solve_gmres(self, matvec_w, tol, restart=20, maxiter=None):
        """scipy GMRES with the reference's arguments: x0 = rhs, restart 20,
        stop at ||r|| <= tol ||b|| (linear_system.py:271-286)."""
        rhs = self.right_hand_side()
        op = scipy.sparse.linalg.LinearOperator((self.n_unk, self.n_unk), dtype=complex,
                                                matvec=lambda v: v - self.apply_t(matvec_w(v)))
        x, info = scipy.sparse.linalg.gmres(op, rhs, x0=rhs, rtol=tol, atol=0.0, restart=restart, maxiter=maxiter)
        return x, info

    def scatter_solution(self, b):
        for i, p in enumerate(self.particles):
            p.scattered_field = b[self.offs[i]:self.offs[i + 1]].copy()


class FarField:
    """Far-field amplitude [2, n_beta, n_alpha] and intensity (postprocessing/far_field.py:9-209)."""

    def __init__(self, polar_angles, azimuthal_angles):
        self.polar_angles = np.asarray(polar_angles)
        self.azimuthal_angles = np.asarray(azimuthal_angles)
        self.amplitude = np.full((2, len(self.polar_angles), len(self.azimuthal_angles)), np.nan, dtype=complex)
        self.signal = np.full((2, len(self.polar_angles), len(self.azimuthal_angles)), np.nan)

    def azimuthal_integral_times_sin_beta(self):
        return np.trapezoid(self.signal, self.azimuthal_angles[None, None, :]) * np.sin(self.polar_angles[None, :])

    def integral(self):
        return np.trapezoid(self.azimuthal_integral_times_sin_beta(), self.polar_angles[None, :])

    def append(self, other):
        if max(self.polar_angles) <= min(other.polar_angles):
            self.polar_angles = np.concatenate((self.polar_angles, other.polar_angles))
            self.signal = np.concatenate((self.signal, other.signal), 1)
            self.amplitude = np.concatenate((self.amplitude, other.amplitude), 1)
        elif min(self.polar_angles) >= max(other.polar_angles):
            self.polar_angles = np.concatenate((other.polar_angles, self.polar_angles))
            self.signal = np.concatenate((other.signal, self.signal), 1)
            self.amplitude = np.concatenate((other.amplitude, self.amplitude), 1)
        else:
            raise ValueError('far fields have overlapping polar angle domains')


def pwe_to_ff_conversion(vacuum_wavelength, pwe):
    """Far field of a PWE (far_field.py:223-259): amplitude -2 pi k kz g, intensity
    2 pi^2 / omega k kz^2 |g|^2, polar angles asin(kappa / k) (pi - .. when downgoing)."""
    omega = fb.angular_frequency(vacuum_wavelength)
    k = pwe.k
    kp = pwe.k_parallel
    polar = np.arcsin(kp / k) if pwe.kind == 'upgoing' else np.pi - np.arcsin(kp / k)
    kz = fb.k_z(k_parallel=kp, k=k)
    ampli = -2 * np.pi * (kz * k)[None, :, None] * pwe.coefficients
    intens = (2 * np.pi ** 2 / omega * (kz ** 2 * k)[None, :, None] * abs(pwe.coefficients) ** 2).real
    idx = np.argsort(polar)
    ff = FarField(np.real(polar[idx]), pwe.azimuthal_angles)
    ff.amplitude = ampli[:, idx, :]
    ff.signal = intens[:, idx, :]
    return ff


def scattered_far_field(vacuum_wavelength, particle_list, layer_system, polar_angles, azimuthal_angles):
    """Scattered far field in the top and bottom layers (far_field.py:363-427)."""
    omega = fb.angular_frequency(vacuum_wavelength)
    polar_angles = np.asarray(polar_angles, dtype=float)
    i_top = layer_system.number_of_layers() - 1
    top = polar_angles[polar_angles < np.pi / 2]
    bot = polar_angles[polar_angles > np.pi / 2]
    n_top = np.sort(np.sin(top) * layer_system.refractive_indices[i_top])
    n_bot = np.sort(np.sin(bot) * layer_system.refractive_indices[0])
    ff_top = ff_bot = None
    if len(top) > 1 and np.imag(layer_system.refractive_indices[i_top]) == 0:
        up, _ = scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, i_top, n_top * omega,
                                    azimuthal_angles)
        ff_top = pwe_to_ff_conversion(vacuum_wavelength, up)
    if len(bot) > 1 and np.imag(layer_system.refractive_indices[0]) == 0:
        _, dn = scattered_field_pwe(vacuum_wavelength, particle_list, layer_system, 0, n_bot * omega,
                                    azimuthal_angles)
        ff_bot = pwe_to_ff_conversion(vacuum_wavelength, dn)
    if ff_top is not None:
        if ff_bot is not None:
            ff_top.append(ff_bot)
        return ff_top
    return ff_bot


def scattering_cross_section(initial_field, particle_list, layer_system, polar_angles, azimuthal_angles):
    """Differential scattering cross section: far-field intensity / |A|^2 n_P / 2 (far_field.py:429-478)."""
    i_top = layer_system.number_of_layers() - 1
    n_p = layer_system.refractive_indices[0 if np.cos(initial_field.polar_angle) > 0 else i_top]
    if np.imag(n_p):
        raise ValueError('plane wave from absorbing layer: cross section undefined')
    i0 = abs(initial_field.amplitude) ** 2 * np.real(n_p) / 2
    ff = scattered_far_field(initial_field.vacuum_wavelength, particle_list, layer_system, polar_angles,
                             azimuthal_angles)
    ff.signal = ff.signal / i0
    return ff


def total_scattering_cross_section(initial_field, particle_list, layer_system, polar_angles, azimuthal_angles):
    """Angle integral of the differential scattering cross section (far_field.py:480-515)."""
    s = scattering_cross_section(initial_field, particle_list, layer_system, polar_angles, azimuthal_angles).integral()
    return s[0] + s[1]
