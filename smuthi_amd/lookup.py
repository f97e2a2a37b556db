"""Host-side assembly of the coefficients the GPU table builder contracts
(prepare_lookup.py:253-464 for the radial table, :16-250 for the volume
tables): direct translation coefficients a5/b5 * P_ld^|dm|, and the
per-kappa layer-mediated coefficients with the trapezoid weights folded in.
Everything here is O(B^2 * N_kappa) and runs once per geometry.
"""
import numpy as np
import scipy.spatial

from . import fields as flds
from . import layers as lay
from . import special as sf


def pair_distance_extremes(xy):
    """(largest, smallest nonzero-index) in-plane distance between particles.

    The largest distance is evaluated on convex-hull vertices with the same
    expression as a full distance matrix, the smallest with a k-d tree."""
    xy = np.asarray(xy, dtype=float)
    n = len(xy)
    if n < 2:
        raise ValueError('a lookup table needs at least two particles')
    cand = np.arange(n)
    if n > 64:
        try:
            cand = scipy.spatial.ConvexHull(xy).vertices
        except scipy.spatial.QhullError:
            cand = np.arange(n)
    sub = xy[cand]
    dmax = 0.0
    for i in range(len(sub)):
        d = np.sqrt((sub[i, 0] - sub[:, 0]) ** 2 + (sub[i, 1] - sub[:, 1]) ** 2)
        dmax = max(dmax, float(d.max()))
    dist, _ = scipy.spatial.cKDTree(xy).query(xy, k=2)
    return dmax, float(dist[:, 1].min())


def radial_grid(rho_max, resolution):
    """The table's distance grid: arange(-3h, rho_max + 3h, h) (prepare_lookup.py:291)."""
    return np.arange(-3 * resolution, rho_max + 3 * resolution, resolution)


def direct_coefficients_radial(l_max, m_max):
    """coef[n1, n2, ld] with direct(rho) = sum_ld coef * h_ld(k rho) for particles at one height."""
    B = flds.blocksize(l_max, m_max)
    tau, ll, mm = flds.index_arrays(l_max, m_max)
    P, _, _ = sf.legendre_normalized(np.array(0.0), np.array(1.0), 2 * l_max)
    coef = np.zeros((B, B, 2 * l_max + 1), complex)
    for n1 in range(B):
        for n2 in range(B):
            l1, m1, l2, m2 = ll[n1], mm[n1], ll[n2], mm[n2]
            dm = abs(m1 - m2)
            for ld in range(max(abs(l1 - l2), dm), l1 + l2 + 1):
                a5, b5 = sf.ab5_coefficients(l2, m2, l1, m1, ld)  # coupling is the transposed translation
                coef[n1, n2, ld] = (a5 if tau[n1] == tau[n2] else b5) * P[ld, dm]
    return coef


def layer_factors(vacuum_wavelength, layer_system, k_parallel, i_s, l_max, m_max):
    """Layer response L[pol, +-, +-, k] and B-dagger / B [pol, +-, n, k] in layer i_s."""
    omega = flds.angular_frequency(vacuum_wavelength)
    k_is = layer_system.wavenumber(i_s, vacuum_wavelength)
    kz = flds.k_z(k_parallel=k_parallel, k=k_is)
    L = np.stack([lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                                  k_parallel, omega, i_s, i_s) for pol in range(2)])
    ct, st = kz / k_is, k_parallel / k_is
    Bd = np.stack([sf.transformation_arrays(l_max, m_max, ct, st, True),
                   sf.transformation_arrays(l_max, m_max, -ct, st, True)], axis=1)
    Bn = np.stack([sf.transformation_arrays(l_max, m_max, ct, st, False),
                   sf.transformation_arrays(l_max, m_max, -ct, st, False)], axis=1)
    return k_is, kz, L, Bd, Bn


def trapezoid_weights(k_parallel):
    """w_k = dk_{k-1} + dk_k, so that sum_k w_k f_k = sum_k (f_k + f_{k+1}) dk_k."""
    dk = np.diff(k_parallel)
    w = np.zeros(len(k_parallel), complex)
    w[:-1] += dk
    w[1:] += dk
    return w


def layer_coefficients_radial(vacuum_wavelength, layer_system, k_parallel, i_s, dz, l_max, m_max):
    """lm[k, n1, n2] with layer-mediated(rho) = sum_k J_|m1-m2|(kappa_k rho) lm[k, n1, n2]."""
    k_is, kz, L, Bd, Bn = layer_factors(vacuum_wavelength, layer_system, k_parallel, i_s, l_max, m_max)
    _, _, mm = flds.index_arrays(l_max, m_max)
    epl, emn = np.exp(2j * kz * dz), np.exp(-2j * kz * dz)
    # belbe[k, n1, n2]
    belbe = np.zeros((len(k_parallel), len(mm), len(mm)), complex)
    for pol in range(2):
        for a, b, ph in ((0, 0, 1.0), (1, 0, emn), (0, 1, epl), (1, 1, 1.0)):
            belbe += (L[pol, a, b] * ph)[:, None, None] * Bd[pol, a].T[:, :, None] * Bn[pol, b].T[:, None, :]
    jac = k_parallel / (kz * k_is)
    dm = np.abs(mm[:, None] - mm[None, :])
    # the table integrand is 2 i^|dm| sum_k (I_k + I_{k+1}) dk_k with I_k = J(kappa_k rho) jac_k belbe_k
    ipow = np.array([1, 1j, -1, -1j])[dm % 4]
    return 2 * ipow[None] * (trapezoid_weights(k_parallel) * jac)[:, None, None] * belbe


def volume_grids(xyz, resolution):
    """rho, z1+z2 and z1-z2 grids of the volume tables (prepare_lookup.py:16-80):
    each arange(min - 3h, max + 3h, h); the two z grids must have one length."""
    xyz = np.asarray(xyz, dtype=float)
    rho_max, rho_min_pair = pair_distance_extremes(xyz[:, :2])
    rho, sz, dz = volume_grids_from_extremes(rho_max, xyz[:, 2].min(), xyz[:, 2].max(), resolution)
    return rho, sz, dz, rho_min_pair


def volume_grids_from_extremes(rho_max, zmin, zmax, resolution):
    """The grids of volume_grids for a largest in-plane distance and a z range
    (a table shared by several ensembles takes the extremes over all of them)."""
    rho = np.arange(-3 * resolution, rho_max + 3 * resolution, resolution)
    sz = np.arange(2 * zmin - 3 * resolution, 2 * zmax + 3 * resolution, resolution)
    dz = np.arange((zmin - zmax) - 3 * resolution, (zmax - zmin) + 3 * resolution, resolution)
    if len(sz) != len(dz):
        raise ValueError('z1+z2 and z1-z2 grids differ in length (%d, %d)' % (len(sz), len(dz)))
    return rho, sz, dz


def direct_coefficients_volume(l_max, m_max):
    """coef[n1, n2, ld] with direct(rho, dz) = sum_ld coef h_ld(k d) P_ld^|m1-m2|(dz / d)."""
    B = flds.blocksize(l_max, m_max)
    tau, ll, mm = flds.index_arrays(l_max, m_max)
    coef = np.zeros((B, B, 2 * l_max + 1), complex)
    for n1 in range(B):
        for n2 in range(B):
            l1, m1, l2, m2 = ll[n1], mm[n1], ll[n2], mm[n2]
            for ld in range(max(abs(l1 - l2), abs(m1 - m2)), l1 + l2 + 1):
                a5, b5 = sf.ab5_coefficients(l2, m2, l1, m1, ld)
                coef[n1, n2, ld] = a5 if tau[n1] == tau[n2] else b5
    return coef


def layer_coefficients_volume(vacuum_wavelength, layer_system, k_parallel, i_s, l_max, m_max):
    """C[k, 4, n1, n2] of the volume tables' layer-mediated parts
    (prepare_lookup.py:138-250): with I(q) = 2 i^|dm| w_k kappa_k / (kz_k k) q,
      plus  = sum_k J (C0 e^{i kz (sz - 2 z_s)} + C1 e^{-i kz (sz - 2 z_s)}),
      minus = sum_k J (C2 e^{i kz dz} + C3 e^{-i kz dz}),
    C0 = I(sum_pol L01 Bd+ B-), C1 = I(sum_pol L10 Bd- B+), C2 = I(sum_pol L00 Bd+ B+), C3 = I(sum_pol L11 Bd- B-)."""
    k_is, kz, L, Bd, Bn = layer_factors(vacuum_wavelength, layer_system, k_parallel, i_s, l_max, m_max)
    _, _, mm = flds.index_arrays(l_max, m_max)
    nk, B = len(k_parallel), len(mm)
    C = np.zeros((nk, 4, B, B), complex)
    for q, (a, b) in enumerate(((0, 1), (1, 0), (0, 0), (1, 1))):
        for pol in range(2):
            C[:, q] += L[pol, a, b][:, None, None] * Bd[pol, a].T[:, :, None] * Bn[pol, b].T[:, None, :]
    jac = k_parallel / (kz * k_is)
    dm = np.abs(mm[:, None] - mm[None, :])
    ipow = np.array([1, 1j, -1, -1j])[dm % 4]
    return C * (2 * ipow[None, None] * (trapezoid_weights(k_parallel) * jac)[:, None, None, None]), kz


def trapezoid_weights_drop_last(k_parallel):
    """Weights of the explicit layer-mediated quadrature: numba_trapz
    (layer_mediated_coupling.py:13-23) sums the intervals 0..n-3 only, so the
    contour's last interval is left out: sum_k w_k f_k = sum_{k<n-2} (f_k + f_{k+1}) dk_k."""
    dk = np.diff(k_parallel)[:-1]
    w = np.zeros(len(k_parallel), complex)
    w[:-2] += dk
    w[1:-1] += dk
    return w


def layer_coefficients_pair(vacuum_wavelength, layer_system, k_parallel, is1, is2, l_max, m_max, weights):
    """C[k, 2a+b, n1, n2] of the explicit layer-mediated block between a
    receiver in layer is1 and an emitter in layer is2
    (layer_mediated_coupling.py:215-340):
      W^R = e^{i (m2-m1) phi} sum_k J_|m1-m2|(kappa_k rho) sum_ab C_ab e^{i s_a kz1 z1} e^{-i s_b kz2 z2},
    s_0 = +1, s_1 = -1, z1 / z2 relative to the layers' reference heights,
    C_ab = 2 i^|m1-m2| w_k kappa_k / (kz2_k k2) sum_pol L_ab(pol) B-dagger_a(pol, n1) B_b(pol, n2)
    with L the response from layer is2 to layer is1 and w the quadrature weights.
    Returns (C, kz1, kz2)."""
    omega = flds.angular_frequency(vacuum_wavelength)
    kp = np.asarray(k_parallel, dtype=complex)
    k1 = complex(layer_system.wavenumber(is1, vacuum_wavelength))
    k2 = complex(layer_system.wavenumber(is2, vacuum_wavelength))
    kz1 = flds.k_z(k_parallel=kp, k=k1)
    kz2 = flds.k_z(k_parallel=kp, k=k2)
    L = np.stack([lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                                  kp, omega, is2, is1) for pol in range(2)])
    Bd = np.stack([sf.transformation_arrays(l_max, m_max, kz1 / k1, kp / k1, True),
                   sf.transformation_arrays(l_max, m_max, -kz1 / k1, kp / k1, True)], axis=1)
    Bn = np.stack([sf.transformation_arrays(l_max, m_max, kz2 / k2, kp / k2, False),
                   sf.transformation_arrays(l_max, m_max, -kz2 / k2, kp / k2, False)], axis=1)
    _, _, mm = flds.index_arrays(l_max, m_max)
    B = len(mm)
    C = np.zeros((len(kp), 4, B, B), complex)
    for a in range(2):
        for b in range(2):
            for pol in range(2):
                C[:, 2 * a + b] += L[pol, a, b][:, None, None] * Bd[pol, a].T[:, :, None] * Bn[pol, b].T[:, None, :]
    jac = kp / (kz2 * k2)
    dm = np.abs(mm[:, None] - mm[None, :])
    ipow = np.array([1, 1j, -1, -1j])[dm % 4]
    C *= 2 * ipow[None, None] * (np.asarray(weights) * jac)[:, None, None, None]
    return C, kz1, kz2
