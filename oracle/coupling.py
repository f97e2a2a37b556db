"""Oracle particle coupling: explicit direct / layer-mediated blocks, the
radial and volume Sommerfeld lookup tables, and the CPU lookup matvecs.

Pinning (tests/test_oracle_coupling.py):
  * direct_coupling_block: W prototype literals of
    tests/unit_tests/multiple_scattering_tests/test_particle_coupling.py:47-68;
    zero self/cross-layer blocks (:67-68).
  * layer_mediated_coupling_block: W^R prototype literals of :22-44 and the
    degenerate-stack identity W^R(3 layers of air) == W(direct) of :71-87.
  * radial_coupling_lookup_table / volumetric_coupling_lookup_table: at every
    grid node the table equals the explicit blocks (direct + layer-mediated,
    with the reference's full trapezoid) for a pair at that separation; the
    rho-mask and the complex64 rounding of the direct part are checked
    directly.
  * radial/volume CPU matvec: equals the dense explicit operator built from
    the same interpolated table entries (brute force on tiny N); linear
    interpolation is exact on linear data; lookup vs explicit operator meets the
    tolerance contracts of test_radial_lookup.py:107-140 and
    test_volumetric_lookup.py:106-139.
"""

import numpy as np
import scipy.interpolate
import scipy.special

from . import basics as fb
from . import layers as lay


def _bessel_jv(n, z):
    return scipy.special.jv(n, z)


def direct_coupling_block(vacuum_wavelength, receiving_particle, emitting_particle, layer_system):
    """Direct coupling block W (linearsystem/particlecoupling/direct_coupling.py:18-105,
    via fields/spherical_harmonic_translations.py:18-64,159-193 with kind
    'outgoing to regular')."""
    lmax1, mmax1 = receiving_particle.l_max, receiving_particle.m_max
    lmax2, mmax2 = emitting_particle.l_max, emitting_particle.m_max
    b1, b2 = fb.blocksize(lmax1, mmax1), fb.blocksize(lmax2, mmax2)
    w = np.zeros((b1, b2), dtype=complex)
    r1, r2 = receiving_particle.position, emitting_particle.position
    i1, i2 = layer_system.layer_number(r1[2]), layer_system.layer_number(r2[2])
    if not ((i1 == i2 or layer_system.is_degenerate()) and emitting_particle is not receiving_particle):
        return w
    k = complex(fb.angular_frequency(vacuum_wavelength) * layer_system.refractive_indices[i1])
    dx, dy, dz = (float(r1[i] - r2[i]) for i in range(3))
    d = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
    phi = np.arctan2(dy, dx)
    legendre = fb.legendre_normalized(dz / d, np.sqrt(dx ** 2 + dy ** 2) / d, lmax1 + lmax2)[0][:, :, 0]
    sph = [complex(fb.spherical_hankel(n, k * d)) for n in range(lmax1 + lmax2 + 1)]
    for m1 in range(-mmax1, mmax1 + 1):
        for m2 in range(-mmax2, mmax2 + 1):
            eimph = np.exp(1j * (m2 - m1) * phi)
            for l1 in range(max(1, abs(m1)), lmax1 + 1):
                for l2 in range(max(1, abs(m2)), lmax2 + 1):
                    A, B = 0j, 0j
                    for ld in range(max(abs(l1 - l2), abs(m1 - m2)), l1 + l2 + 1):
                        a5, b5 = fb.ab5_coefficients(l2, m2, l1, m1, ld)
                        A += a5 * sph[ld] * legendre[ld, abs(m1 - m2)]
                        B += b5 * sph[ld] * legendre[ld, abs(m1 - m2)]
                    A, B = eimph * A, eimph * B
                    for tau1 in range(2):
                        n1 = fb.multi_to_single_index(tau1, l1, m1, lmax1, mmax1)
                        for tau2 in range(2):
                            n2 = fb.multi_to_single_index(tau2, l2, m2, lmax2, mmax2)
                            w[n1, n2] = A if tau1 == tau2 else B
    return w


def _trapz_drop_last(y, x):
    """The reference's numba_trapz sums intervals 0..len-3, i.e. it drops the
    last interval (layer_mediated_coupling.py:13-23)."""
    out = 0j
    for i in range(len(y) - 2):
        out += (x[i + 1] - x[i]) * (y[i + 1] + y[i]) / 2.0
    return out


def layer_mediated_coupling_block(vacuum_wavelength, receiving_particle, emitting_particle, layer_system,
                                  k_parallel, full_trapezoid=False):
    """Explicit layer-mediated block W^R (layer_mediated_coupling.py:215-340).

    ``full_trapezoid=True`` integrates over every contour interval, which is
    what the lookup tables do; the default reproduces the reference's
    dropped last interval."""
    omega = fb.angular_frequency(vacuum_wavelength)
    lmax1, mmax1 = receiving_particle.l_max, receiving_particle.m_max
    lmax2, mmax2 = emitting_particle.l_max, emitting_particle.m_max
    b1, b2 = fb.blocksize(lmax1, mmax1), fb.blocksize(lmax2, mmax2)
    rs1 = np.array(receiving_particle.position, dtype=float)
    rs2 = np.array(emitting_particle.position, dtype=float)
    rs2s1 = rs1 - rs2
    rho = np.linalg.norm(rs2s1[0:2])
    phi = np.arctan2(rs2s1[1], rs2s1[0])
    is1 = layer_system.layer_number(rs1[2])
    z1 = rs1[2] - layer_system.reference_z(is1)
    is2 = layer_system.layer_number(rs2[2])
    z2 = rs2[2] - layer_system.reference_z(is2)
    k1 = omega * layer_system.refractive_indices[is1]
    k2 = omega * layer_system.refractive_indices[is2]
    kz1 = fb.k_z(k_parallel=k_parallel, k=k1)
    kz2 = fb.k_z(k_parallel=k_parallel, k=k2)
    ejkz = np.zeros((2, 2, len(k_parallel)), dtype=complex)
    ejkz[0, 0] = np.exp(1j * kz1 * z1)
    ejkz[0, 1] = np.exp(-1j * kz1 * z1)
    ejkz[1, 0] = np.exp(1j * kz2 * z2)
    ejkz[1, 1] = np.exp(-1j * kz2 * z2)
    L = np.zeros((2, 2, 2, len(k_parallel)), dtype=complex)
    for pol in range(2):
        L[pol] = lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                                 k_parallel, omega, is2, is1)
    B = [np.zeros((2, 2, b1, len(k_parallel)), dtype=complex), np.zeros((2, 2, b2, len(k_parallel)), dtype=complex)]
    m_vec = [np.zeros(b1, dtype=int), np.zeros(b2, dtype=int)]
    for ip, (kk, kzz, lm, mm, dag) in enumerate(((k1, kz1, lmax1, mmax1, True), (k2, kz2, lmax2, mmax2, False))):
        _, pi_pl, tau_pl = fb.legendre_normalized(kzz / kk, k_parallel / kk, lm)
        _, pi_mn, tau_mn = fb.legendre_normalized(-kzz / kk, k_parallel / kk, lm)
        pis, taus = (pi_pl, pi_mn), (tau_pl, tau_mn)
        for n, tau, l, m in fb.index_specs(lm, mm):
            m_vec[ip][n] = m
            for iplmn in range(2):
                for pol in range(2):
                    B[ip][pol, iplmn, n] = fb.transformation_coefficients_vwf(
                        tau, l, m, pol, pilm_list=pis[iplmn], taulm_list=taus[iplmn], dagger=dag)
    bessel = [_bessel_jv(dm, k_parallel * rho) for dm in range(lmax1 + lmax2 + 1)]
    jacobi = k_parallel / (kz2 * k2)
    m2m1 = m_vec[1][None, :] - m_vec[0][:, None]
    wr_const = 4 * (1j) ** abs(m2m1) * np.exp(1j * m2m1 * phi)
    integral = np.zeros((b1, b2), dtype=complex)
    for n1 in range(b1):
        BeL = np.zeros((2, 2, len(k_parallel)), dtype=complex)
        for iplmn1 in range(2):
            for pol in range(2):
                BeL[pol] += L[pol, iplmn1] * B[0][pol, iplmn1, n1] * ejkz[0, iplmn1]
        for n2 in range(b2):
            BeLBe = np.zeros(len(k_parallel), dtype=complex)
            for iplmn2 in range(2):
                for pol in range(2):
                    BeLBe += BeL[pol, iplmn2] * B[1][pol, iplmn2, n2] * ejkz[1, 1 - iplmn2]
            integrand = bessel[abs(m_vec[0][n1] - m_vec[1][n2])] * jacobi * BeLBe
            if full_trapezoid:
                integral[n1, n2] = np.sum((integrand[1:] + integrand[:-1]) * np.diff(k_parallel)) / 2
            else:
                integral[n1, n2] = _trapz_drop_last(integrand, k_parallel)
    return wr_const * integral


def _m_list(l_max, m_max):
    ml = [0] * fb.blocksize(l_max, m_max)
    for n, tau, l, m in fb.index_specs(l_max, m_max):
        ml[n] = m
    return ml


def _layer_mediated_factors(vacuum_wavelength, layer_system, k_parallel, i_s, l_max, m_max):
    """Layer response and B / B-dagger arrays shared by both tables
    (prepare_lookup.py:138-167 and :360-388)."""
    blocksize = fb.blocksize(l_max, m_max)
    k_is = layer_system.wavenumber(i_s, vacuum_wavelength)
    kz_is = fb.k_z(k_parallel=k_parallel, k=k_is)
    L = np.zeros((2, 2, 2, len(k_parallel)), dtype=complex)
    for pol in range(2):
        L[pol] = lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                                 k_parallel, fb.angular_frequency(vacuum_wavelength), i_s, i_s)
    B_dag = np.zeros((2, 2, blocksize, len(k_parallel)), dtype=complex)
    B = np.zeros((2, 2, blocksize, len(k_parallel)), dtype=complex)
    _, pi_pl, tau_pl = fb.legendre_normalized(kz_is / k_is, k_parallel / k_is, l_max)
    _, pi_mn, tau_mn = fb.legendre_normalized(-kz_is / k_is, k_parallel / k_is, l_max)
    for n, tau, l, m in fb.index_specs(l_max, m_max):
        for pol in range(2):
            B_dag[pol, 0, n] = fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_pl, taulm_list=tau_pl,
                                                                  dagger=True)
            B_dag[pol, 1, n] = fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_mn, taulm_list=tau_mn,
                                                                  dagger=True)
            B[pol, 0, n] = fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_pl, taulm_list=tau_pl)
            B[pol, 1, n] = fb.transformation_coefficients_vwf(tau, l, m, pol, pilm_list=pi_mn, taulm_list=tau_mn)
    return k_is, kz_is, L, B_dag, B


def radial_coupling_lookup_table(vacuum_wavelength, particle_list, layer_system, k_parallel, resolution=None):
    """Radial lookup table [len_rho, B, B] and its rho grid
    (linearsystem/particlecoupling/prepare_lookup.py:253-464, CPU branch).

    The direct part is stored complex64 then added to the complex128
    layer-mediated part, as in the reference (:301, :459)."""
    if resolution is None:
        resolution = vacuum_wavelength / 100
    l_max = max(p.l_max for p in particle_list)
    m_max = max(p.m_max for p in particle_list)
    bs = fb.blocksize(l_max, m_max)
    x = np.array([p.position[0] for p in particle_list], dtype=float)
    y = np.array([p.position[1] for p in particle_list], dtype=float)
    rho_pp = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
    rho = np.arange(-3 * resolution, rho_pp.max() + 3 * resolution, resolution)
    z = particle_list[0].position[2]
    i_s = layer_system.layer_number(z)
    k_is = layer_system.wavenumber(i_s, vacuum_wavelength)
    dz = z - layer_system.reference_z(i_s)
    len_rho = len(rho)
    w = np.zeros((len_rho, bs, bs), dtype=np.complex64)
    bessel_h = []
    with np.errstate(invalid='ignore'):
        for n in range(2 * l_max + 1):
            h = fb.spherical_hankel(n, k_is * rho)
            h[rho <= 0] = np.nan
            bessel_h.append(h)
    legendre, _, _ = fb.legendre_normalized(np.array([0.0]), np.array([1.0]), 2 * l_max)
    ml = _m_list(l_max, m_max)
    for m1 in range(-m_max, m_max + 1):
        for m2 in range(-m_max, m_max + 1):
            for l1 in range(max(1, abs(m1)), l_max + 1):
                for l2 in range(max(1, abs(m2)), l_max + 1):
                    A = np.zeros(len_rho, dtype=complex)
                    Bv = np.zeros(len_rho, dtype=complex)
                    for ld in range(max(abs(l1 - l2), abs(m1 - m2)), l1 + l2 + 1):
                        a5, b5 = fb.ab5_coefficients(l2, m2, l1, m1, ld)
                        A = A + a5 * bessel_h[ld] * legendre[ld, abs(m1 - m2)]
                        Bv = Bv + b5 * bessel_h[ld] * legendre[ld, abs(m1 - m2)]
                    for tau1 in range(2):
                        n1 = fb.multi_to_single_index(tau1, l1, m1, l_max, m_max)
                        for tau2 in range(2):
                            n2 = fb.multi_to_single_index(tau2, l2, m2, l_max, m_max)
                            w[:, n1, n2] = A if tau1 == tau2 else Bv
    offdiag = rho_pp[~np.eye(rho_pp.shape[0], dtype=bool)]
    w[rho < offdiag.min() / 2] = 0
    if not layer_system.is_degenerate():
        k_is, kz_is, L, B_dag, B = _layer_mediated_factors(vacuum_wavelength, layer_system, k_parallel, i_s,
                                                            l_max, m_max)
        epl2 = np.exp(2j * kz_is * dz)
        emn2 = np.exp(-2j * kz_is * dz)
        wr = np.zeros((len_rho, bs, bs), dtype=complex)
        dkp = np.diff(k_parallel)
        for dm in range(2 * m_max + 1):
            besjac = _bessel_jv(dm, k_parallel[None, :] * rho[:, None]) * (k_parallel / (kz_is * k_is))[None, :]
            for n1 in range(bs):
                for n2 in range(bs):
                    if abs(ml[n1] - ml[n2]) != dm:
                        continue
                    belbe = np.zeros(len(k_parallel), dtype=complex)
                    for pol in range(2):
                        belbe += L[pol, 0, 0] * B_dag[pol, 0, n1] * B[pol, 0, n2]
                        belbe += L[pol, 1, 0] * B_dag[pol, 1, n1] * B[pol, 0, n2] * emn2
                        belbe += L[pol, 0, 1] * B_dag[pol, 0, n1] * B[pol, 1, n2] * epl2
                        belbe += L[pol, 1, 1] * B_dag[pol, 1, n1] * B[pol, 1, n2]
                    integrand = besjac * belbe[None, :]
                    wr[:, n1, n2] = 2 * (1j) ** abs(ml[n2] - ml[n1]) * (
                        (integrand[:, :-1] + integrand[:, 1:]) * dkp[None, :]).sum(axis=-1)
        w = w + wr
    return w, rho


def volumetric_coupling_lookup_table(vacuum_wavelength, particle_list, layer_system, k_parallel, resolution=None):
    """Volume lookup tables w_pl[rho, sz, B, B], w_mn[rho, dz, B, B] (complex64)
    and grids (linearsystem/particlecoupling/prepare_lookup.py:16-250, CPU branch)."""
    if resolution is None:
        resolution = vacuum_wavelength / 100
    l_max = max(p.l_max for p in particle_list)
    m_max = max(p.m_max for p in particle_list)
    bs = fb.blocksize(l_max, m_max)
    x = np.array([p.position[0] for p in particle_list], dtype=float)
    y = np.array([p.position[1] for p in particle_list], dtype=float)
    zz = np.array([p.position[2] for p in particle_list], dtype=float)
    rho_pp = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
    dz_min, dz_max = zz.min() - zz.max(), zz.max() - zz.min()
    sz_min, sz_max = 2 * zz.min(), 2 * zz.max()
    rho = np.arange(-3 * resolution, rho_pp.max() + 3 * resolution, resolution)
    sz = np.arange(sz_min - 3 * resolution, sz_max + 3 * resolution, resolution)
    dz = np.arange(dz_min - 3 * resolution, dz_max + 3 * resolution, resolution)
    assert len(sz) == len(dz)
    i_s = layer_system.layer_number(particle_list[0].position[2])
    k_is = layer_system.wavenumber(i_s, vacuum_wavelength)
    z_is = layer_system.reference_z(i_s)
    len_rho, len_z = len(rho), len(dz)
    wr_pl = np.zeros((len_rho, len_z, bs, bs), dtype=np.complex64)
    wr_mn = np.zeros((len_rho, len_z, bs, bs), dtype=np.complex64)
    w = np.zeros((len_rho, len_z, bs, bs), dtype=np.complex64)
    r = np.sqrt(dz[None, :] ** 2 + rho[:, None] ** 2)
    r[r == 0] = 1e-20
    legendre, _, _ = fb.legendre_normalized(dz[None, :] / r, rho[:, None] / r, 2 * l_max)
    with np.errstate(invalid='ignore'):
        bessel_h = [fb.spherical_hankel(dm, k_is * r) for dm in range(2 * l_max + 1)]
    ml = _m_list(l_max, m_max)
    for m1 in range(-m_max, m_max + 1):
        for m2 in range(-m_max, m_max + 1):
            for l1 in range(max(1, abs(m1)), l_max + 1):
                for l2 in range(max(1, abs(m2)), l_max + 1):
                    A = np.zeros((len_rho, len_z), dtype=complex)
                    Bv = np.zeros((len_rho, len_z), dtype=complex)
                    for ld in range(max(abs(l1 - l2), abs(m1 - m2)), l1 + l2 + 1):
                        a5, b5 = fb.ab5_coefficients(l2, m2, l1, m1, ld)
                        A += a5 * bessel_h[ld] * legendre[ld, abs(m1 - m2)]
                        Bv += b5 * bessel_h[ld] * legendre[ld, abs(m1 - m2)]
                    for tau1 in range(2):
                        n1 = fb.multi_to_single_index(tau1, l1, m1, l_max, m_max)
                        for tau2 in range(2):
                            n2 = fb.multi_to_single_index(tau2, l2, m2, l_max, m_max)
                            w[:, :, n1, n2] = A if tau1 == tau2 else Bv
    offdiag = rho_pp[~np.eye(rho_pp.shape[0], dtype=bool)]
    w[rho < offdiag.min() / 2, :, :, :] = 0
    if not layer_system.is_degenerate():
        k_is, kz_is, L, B_dag, B = _layer_mediated_factors(vacuum_wavelength, layer_system, k_parallel, i_s,
                                                            l_max, m_max)
        epl_sz = np.exp(1j * kz_is[None, :] * (sz[:, None] - 2 * z_is))
        emn_sz = np.exp(-1j * kz_is[None, :] * (sz[:, None] - 2 * z_is))
        epl_dz = np.exp(1j * kz_is[None, :] * dz[:, None])
        emn_dz = np.exp(-1j * kz_is[None, :] * dz[:, None])
        dkp = np.diff(k_parallel)
        for dm in range(2 * m_max + 1):
            besjac = _bessel_jv(dm, k_parallel[None, :] * rho[:, None]) * (k_parallel / (kz_is * k_is))[None, :]
            for n1 in range(bs):
                for n2 in range(bs):
                    if abs(ml[n1] - ml[n2]) != dm:
                        continue
                    bpl = np.zeros((len_z, len(k_parallel)), dtype=complex)
                    bmn = np.zeros((len_z, len(k_parallel)), dtype=complex)
                    for pol in range(2):
                        bpl += ((L[pol, 0, 1] * B_dag[pol, 0, n1] * B[pol, 1, n2])[None, :] * epl_sz
                                + (L[pol, 1, 0] * B_dag[pol, 1, n1] * B[pol, 0, n2])[None, :] * emn_sz)
                        bmn += ((L[pol, 0, 0] * B_dag[pol, 0, n1] * B[pol, 0, n2])[None, :] * epl_dz
                                + (L[pol, 1, 1] * B_dag[pol, 1, n1] * B[pol, 1, n2])[None, :] * emn_dz)
                    ph = 2 * (1j) ** abs(ml[n2] - ml[n1])
                    integ = besjac[:, None, :] * bpl[None, :, :]
                    wr_pl[:, :, n1, n2] = ph * ((integ[:, :, :-1] + integ[:, :, 1:]) * dkp[None, None, :]).sum(-1)
                    integ = besjac[:, None, :] * bmn[None, :, :]
                    wr_mn[:, :, n1, n2] = ph * ((integ[:, :, :-1] + integ[:, :, 1:]) * dkp[None, None, :]).sum(-1)
    return wr_pl, w + wr_mn, rho, sz, dz


def system_index_maps(particle_list):
    """Per unknown: particle number, global lookup index n, order m, and the
    offsets of each particle's block (linear_system.py:350-385, :877-900)."""
    l_max = max(p.l_max for p in particle_list)
    m_max = max(p.m_max for p in particle_list)
    part, nlk, mm, offs = [], [], [], []
    off = 0
    for i, p in enumerate(particle_list):
        offs.append(off)
        for n, tau, l, m in fb.index_specs(p.l_max, p.m_max):
            part.append(i)
            nlk.append(fb.multi_to_single_index(tau, l, m, l_max, m_max))
            mm.append(m)
        off += fb.blocksize(p.l_max, p.m_max)
    return np.array(part), np.array(nlk), np.array(mm), np.array(offs + [off])


def catmull_rom(w, ym1, y0, y1, y2):
    """4-tap cubic Hermite (Catmull-Rom) interpolant used by the reference's
    GPU kernels (linear_system_cuda.py:348-353)."""
    return ((-w ** 3 + 2 * w ** 2 - w) * ym1 + (3 * w ** 3 - 5 * w ** 2 + 2) * y0
            + (-3 * w ** 3 + 4 * w ** 2 + w) * y1 + (w ** 3 - w ** 2) * y2) / 2


def radial_interpolated_entries(table, rho_grid, rho, kind):
    """Interpolated table entries at separations ``rho`` (array), all (n1,n2):
    returns [len(rho), B, B]. 'linear' and 'cubic' follow scipy interp1d
    (CPU path, linear_system.py:954-956); 'catmull-rom' follows the GPU
    kernel (linear_system_cuda.py:342-405)."""
    rho = np.asarray(rho, dtype=float)
    if kind in ('linear', 'cubic'):
        f = scipy.interpolate.interp1d(rho_grid, table, kind=kind, axis=0, assume_sorted=True)
        return f(rho)
    h = rho_grid[1] - rho_grid[0]
    t = (rho - rho_grid[0]) / h
    r = np.floor(t).astype(int)
    w = (t - r)[:, None, None]
    return catmull_rom(w, table[r - 1], table[r], table[r + 1], table[r + 2])


class RadialLookupCPU:
    """CPU radial-lookup coupling operator (linear_system.py:903-975).

    The reference loops over (n1, n2) with per-pair interp1d objects; this
    restatement interpolates all (n1, n2) of one particle pair at once, which
    evaluates the same interpolant on the same data. ``kind`` may also be
    'catmull-rom' for the GPU-kernel interpolant."""

    def __init__(self, table, rho_grid, particle_list, kind='linear'):
        self.table = table
        self.rho_grid = rho_grid
        self.kind = kind
        self.part, self.nlk, self.m, self.offs = system_index_maps(particle_list)
        x = np.array([p.position[0] for p in particle_list], dtype=float)
        y = np.array([p.position[1] for p in particle_list], dtype=float)
        self.rho = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
        self.phi = np.arctan2(y[:, None] - y[None, :], x[:, None] - x[None, :])
        self.n_unk = len(self.part)
        self.n_part = len(particle_list)

    def block(self, i1, i2):
        """Coupling block between particles i1 (receiving) and i2 (emitting)."""
        s1 = slice(self.offs[i1], self.offs[i1 + 1])
        s2 = slice(self.offs[i2], self.offs[i2 + 1])
        ent = radial_interpolated_entries(self.table, self.rho_grid, np.array([self.rho[i1, i2]]), self.kind)[0]
        blk = ent[np.ix_(self.nlk[s1], self.nlk[s2])]
        return blk * np.exp(1j * (self.m[s2][None, :] - self.m[s1][:, None]) * self.phi[i1, i2])

    def matvec(self, v):
        v = np.asarray(v)
        out = np.zeros(v.shape, dtype=complex)
        np_ = self.n_part
        ii1, ii2 = np.meshgrid(np.arange(np_), np.arange(np_), indexing='ij')
        ent = radial_interpolated_entries(self.table, self.rho_grid, self.rho.ravel(), self.kind)
        ent = ent.reshape(np_, np_, ent.shape[1], ent.shape[2])
        for i1 in range(np_):
            s1 = slice(self.offs[i1], self.offs[i1 + 1])
            for i2 in range(np_):
                s2 = slice(self.offs[i2], self.offs[i2 + 1])
                blk = ent[i1, i2][np.ix_(self.nlk[s1], self.nlk[s2])]
                ph = np.exp(1j * (self.m[s2][None, :] - self.m[s1][:, None]) * self.phi[i1, i2])
                out[s1] += (blk * ph) @ v[s2]
        return out

    def dense(self):
        W = np.zeros((self.n_unk, self.n_unk), dtype=complex)
        np_ = self.n_part
        ent = radial_interpolated_entries(self.table, self.rho_grid, self.rho.ravel(), self.kind)
        ent = ent.reshape(np_, np_, ent.shape[1], ent.shape[2])
        for i1 in range(np_):
            s1 = slice(self.offs[i1], self.offs[i1 + 1])
            for i2 in range(np_):
                s2 = slice(self.offs[i2], self.offs[i2 + 1])
                blk = ent[i1, i2][np.ix_(self.nlk[s1], self.nlk[s2])]
                W[s1, s2] = blk * np.exp(1j * (self.m[s2][None, :] - self.m[s1][:, None]) * self.phi[i1, i2])
        return W


class RadialLookupCPUReferenceLoop:
    """Literal restatement of the reference's per-(n1,n2) matvec loop with
    scipy interp1d objects (linear_system.py:903-975). Used as the timed CPU
    baseline and to cross-check RadialLookupCPU."""

    def __init__(self, table, rho_grid, particle_list, kind='linear', n1_rows=None):
        l_max = max(p.l_max for p in particle_list)
        m_max = max(p.m_max for p in particle_list)
        self.bs = fb.blocksize(l_max, m_max)
        x = np.array([p.position[0] for p in particle_list], dtype=float)
        y = np.array([p.position[1] for p in particle_list], dtype=float)
        self.rho = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
        self.phi = np.arctan2(y[:, None] - y[None, :], x[:, None] - x[None, :])
        self.sys_idx = [[] for _ in range(self.bs)]
        self.pnum = [[] for _ in range(self.bs)]
        self.m_list = [None] * self.bs
        off = 0
        for i, p in enumerate(particle_list):
            for n, tau, l, m in fb.index_specs(p.l_max, p.m_max):
                nl = fb.multi_to_single_index(tau, l, m, l_max, m_max)
                self.sys_idx[nl].append(off + n)
                self.pnum[nl].append(i)
                self.m_list[nl] = m
            off += fb.blocksize(p.l_max, p.m_max)
        self.sys_idx = [np.array(a, dtype=int) for a in self.sys_idx]
        self.pnum = [np.array(a, dtype=int) for a in self.pnum]
        # (interpolants only for the rows n1_rows when given: a worker of a parallel timing)
        rows = range(self.bs) if n1_rows is None else set(n1_rows)
        self.lookup = [[scipy.interpolate.interp1d(rho_grid, table[:, n1, n2], kind=kind, axis=-1,
                                                   assume_sorted=True) for n2 in range(self.bs)]
                       if n1 in rows else None for n1 in range(self.bs)]

    def matvec(self, v, n1_subset=None):
        out = np.zeros(v.shape, dtype=complex)
        rows = range(self.bs) if n1_subset is None else n1_subset
        for n1 in rows:
            i1, idx1, m1 = self.pnum[n1], self.sys_idx[n1], self.m_list[n1]
            for n2 in range(self.bs):
                i2, idx2, m2 = self.pnum[n2], self.sys_idx[n2], self.m_list[n2]
                rho = self.rho[i1[:, None], i2[None, :]]
                phi = self.phi[i1[:, None], i2[None, :]]
                M = self.lookup[n1][n2](rho) * np.exp(1j * (m2 - m1) * phi)
                out[idx1] += M.dot(v[idx2])
        return out


class VolumeLookupCPU:
    """CPU volume-lookup coupling operator (linear_system.py:593-685):
    per (n1,n2) RectBivariateSpline of order 1 or 3 over (rho, z1+z2) and
    (rho, z1-z2); also 'catmull-rom' for the GPU-kernel tensor interpolant
    (linear_system_cuda.py:149-256)."""

    def __init__(self, table_pl, table_mn, rho_grid, sz_grid, dz_grid, particle_list, kind='linear'):
        self.tpl, self.tmn = table_pl, table_mn
        self.rho_grid, self.sz_grid, self.dz_grid = rho_grid, sz_grid, dz_grid
        self.kind = kind
        self.part, self.nlk, self.m, self.offs = system_index_maps(particle_list)
        x = np.array([p.position[0] for p in particle_list], dtype=float)
        y = np.array([p.position[1] for p in particle_list], dtype=float)
        z = np.array([p.position[2] for p in particle_list], dtype=float)
        self.rho = np.sqrt((x[:, None] - x[None, :]) ** 2 + (y[:, None] - y[None, :]) ** 2)
        self.phi = np.arctan2(y[:, None] - y[None, :], x[:, None] - x[None, :])
        self.sz = z[:, None] + z[None, :]
        self.dz = z[:, None] - z[None, :]
        self.n_part = len(particle_list)
        self.n_unk = len(self.part)
        self.bs = table_pl.shape[-1]
        if kind in ('linear', 'cubic'):
            order = 3 if kind == 'cubic' else 1
            self.spl = {}
            for n1 in range(self.bs):
                for n2 in range(self.bs):
                    self.spl[n1, n2] = [
                        scipy.interpolate.RectBivariateSpline(rho_grid, sz_grid, table_pl[:, :, n1, n2].real,
                                                              kx=order, ky=order),
                        scipy.interpolate.RectBivariateSpline(rho_grid, sz_grid, table_pl[:, :, n1, n2].imag,
                                                              kx=order, ky=order),
                        scipy.interpolate.RectBivariateSpline(rho_grid, dz_grid, table_mn[:, :, n1, n2].real,
                                                              kx=order, ky=order),
                        scipy.interpolate.RectBivariateSpline(rho_grid, dz_grid, table_mn[:, :, n1, n2].imag,
                                                              kx=order, ky=order)]

    def entries(self, rho, sz, dz):
        """Interpolated Lpl+Lmn for arrays of (rho, sz, dz): [K, B, B]."""
        rho, sz, dz = (np.atleast_1d(np.asarray(a, dtype=float)) for a in (rho, sz, dz))
        out = np.zeros((len(rho), self.bs, self.bs), dtype=complex)
        if self.kind in ('linear', 'cubic'):
            for (n1, n2), s in self.spl.items():
                out[:, n1, n2] = (s[0].ev(rho, sz) + 1j * s[1].ev(rho, sz)
                                  + s[2].ev(rho, dz) + 1j * s[3].ev(rho, dz))
            return out
        h = self.rho_grid[1] - self.rho_grid[0]

        def lookup2d(tab, zmin, zval):
            tr = (rho - self.rho_grid[0]) / h
            ir = np.floor(tr).astype(int)
            wr = tr - ir
            tz = (zval - zmin) / h
            iz = np.floor(tz).astype(int)
            wz = tz - iz
            rows = []
            for dr in (-1, 0, 1, 2):
                cols = [tab[ir + dr, iz + dzz] for dzz in (-1, 0, 1, 2)]
                rows.append(catmull_rom(wz[:, None, None], *cols))
            return catmull_rom(wr[:, None, None], *rows)

        return lookup2d(self.tpl, self.sz_grid[0], sz) + lookup2d(self.tmn, self.dz_grid[0], dz)

    def block(self, i1, i2):
        s1 = slice(self.offs[i1], self.offs[i1 + 1])
        s2 = slice(self.offs[i2], self.offs[i2 + 1])
        ent = self.entries(self.rho[i1, i2], self.sz[i1, i2], self.dz[i1, i2])[0]
        blk = ent[np.ix_(self.nlk[s1], self.nlk[s2])]
        return blk * np.exp(1j * (self.m[s2][None, :] - self.m[s1][:, None]) * self.phi[i1, i2])

    def dense(self):
        W = np.zeros((self.n_unk, self.n_unk), dtype=complex)
        ent = self.entries(self.rho.ravel(), self.sz.ravel(), self.dz.ravel())
        ent = ent.reshape(self.n_part, self.n_part, self.bs, self.bs)
        for i1 in range(self.n_part):
            s1 = slice(self.offs[i1], self.offs[i1 + 1])
            for i2 in range(self.n_part):
                s2 = slice(self.offs[i2], self.offs[i2 + 1])
                blk = ent[i1, i2][np.ix_(self.nlk[s1], self.nlk[s2])]
                W[s1, s2] = blk * np.exp(1j * (self.m[s2][None, :] - self.m[s1][:, None]) * self.phi[i1, i2])
        return W

    def matvec(self, v):
        return self.dense() @ np.asarray(v)
