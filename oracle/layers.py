"""Oracle planar layer system: geometry helpers and the S-/T-matrix layer
response (layers.py:17-380), plus sphere Mie coefficients (t_matrix.py:8-93).

Pinning (tests/test_oracle_basics.py):
  * layersystem_response_matrix: TSPL prototype literals of
    tests/unit_tests/test_layerresponse.py:43-52; S-matrix == T-matrix identity
    of :37-41; mpmath at 30 digits reproduces the numpy path (:18-35).
  * mie_coefficient: prototype literals of t_matrix_tests/test_t_matrix.py:23-27;
    closed form: a sphere with the medium's index has zero Mie coefficients.
"""

import numpy as np

from . import basics as fb


class LayerSystem:
    """Planar stack, z=0 at the interface between layers 0 and 1 (layers.py:17-202)."""

    def __init__(self, thicknesses=None, refractive_indices=None):
        thicknesses = [0, 0] if thicknesses is None else list(thicknesses)
        refractive_indices = [1, 1] if refractive_indices is None else list(refractive_indices)
        self.thicknesses = list(thicknesses)
        self.thicknesses[0] = 0
        self.thicknesses[-1] = 0
        self.refractive_indices = list(refractive_indices)

    def is_degenerate(self):
        return self.thicknesses == [0, 0] and self.refractive_indices[0] == self.refractive_indices[-1]

    def number_of_layers(self):
        return len(self.thicknesses)

    def lower_zlimit(self, i):
        if i == 0:
            return -np.inf
        return float(sum(self.thicknesses[1:i]))

    def upper_zlimit(self, i):
        if i == self.number_of_layers() - 1:
            return np.inf
        return float(sum(self.thicknesses[1:i + 1]))

    def reference_z(self, i):
        return self.upper_zlimit(i) if i == 0 else self.lower_zlimit(i)

    def layer_number(self, z):
        d = 0
        laynum = 0
        for th in self.thicknesses[1:]:
            if z >= d:
                laynum += 1
                d += th
            else:
                break
        return laynum

    def wavenumber(self, layer_number, vacuum_wavelength):
        return self.refractive_indices[layer_number] * fb.angular_frequency(vacuum_wavelength)


def fresnel_r(pol, kz1, kz2, n1, n2):
    if pol == 0:
        return (kz1 - kz2) / (kz1 + kz2)
    return (n2 ** 2 * kz1 - n1 ** 2 * kz2) / (n2 ** 2 * kz1 + n1 ** 2 * kz2)


def fresnel_t(pol, kz1, kz2, n1, n2):
    if pol == 0:
        return 2 * kz1 / (kz1 + kz2)
    return 2 * n1 * n2 * kz1 / (n2 ** 2 * kz1 + n1 ** 2 * kz2)


def _kz_list(layer_n, kpar, omega, mp=None):
    out = []
    for n in layer_n:
        if mp is None:
            kz = np.sqrt((omega * n) ** 2 - kpar ** 2 + 0j)
        else:
            kz = mp.sqrt((omega * n) ** 2 - kpar ** 2 + 0j)
        if kz.imag < 0:
            kz = -kz
        out.append(kz)
    return out


def _mat(a, mp):
    return mp.matrix(a) if mp is not None else np.array(a, dtype=complex)


def _mul(a, b, mp):
    return a * b if mp is not None else a @ b


def _inv(a, mp):
    return a ** -1 if mp is not None else np.linalg.inv(a)


def _exp(x, mp):
    return mp.exp(x) if mp is not None else np.exp(x)


def layersystem_transfer_matrix(pol, layer_d, layer_n, kpar, omega, mp=None):
    """Transfer matrix (layers.py:270-294)."""
    kzs = _kz_list(layer_n, kpar, omega, mp)
    tmat = mp.eye(2) if mp is not None else np.eye(2, dtype=complex)
    for i in range(len(layer_d) - 1):
        t = fresnel_t(pol, kzs[i], kzs[i + 1], layer_n[i], layer_n[i + 1])
        r = fresnel_r(pol, kzs[i], kzs[i + 1], layer_n[i], layer_n[i + 1])
        dmat = _mat([[1 / t, r / t], [r / t, 1 / t]], mp)
        pmat = _mat([[_exp(-1j * kzs[i] * layer_d[i], mp), 0], [0, _exp(1j * kzs[i] * layer_d[i], mp)]], mp)
        tmat = _mul(tmat, _mul(pmat, dmat, mp), mp)
    return tmat


def layersystem_scattering_matrix(pol, layer_d, layer_n, kpar, omega, mp=None):
    """Scattering matrix (layers.py:297-327)."""
    kzs = _kz_list(layer_n, kpar, omega, mp)
    smat = mp.eye(2) if mp is not None else np.eye(2, dtype=complex)
    for i in range(len(layer_d) - 1):
        t = fresnel_t(pol, kzs[i], kzs[i + 1], layer_n[i], layer_n[i + 1])
        r = fresnel_r(pol, kzs[i], kzs[i + 1], layer_n[i], layer_n[i + 1])
        dmat = _mat([[1 / t, r / t], [r / t, 1 / t]], mp)
        pmat = _mat([[_exp(-1j * kzs[i] * layer_d[i], mp), 0], [0, _exp(1j * kzs[i] * layer_d[i], mp)]], mp)
        tmat = _mul(pmat, dmat, mp)
        s11 = smat[0, 0] / (tmat[0, 0] - smat[0, 1] * tmat[1, 0])
        s12 = (smat[0, 1] * tmat[1, 1] - tmat[0, 1]) / (tmat[0, 0] - smat[0, 1] * tmat[1, 0])
        s21 = smat[1, 1] * tmat[1, 0] * s11 + smat[1, 0]
        s22 = smat[1, 1] * tmat[1, 0] * s12 + smat[1, 1] * tmat[1, 1]
        smat = _mat([[s11, s12], [s21, s22]], mp)
    return smat


def layersystem_response_matrix(pol, layer_d, layer_n, kpar, omega, fromlayer, tolayer, mp=None):
    """2x2 layer response L[+-,+-] (layers.py:330-380); array kpar gives [2,2,N]."""
    if hasattr(kpar, '__len__'):
        res = np.zeros((2, 2, len(kpar)), dtype=complex)
        for i, kp in enumerate(kpar):
            res[:, :, i] = layersystem_response_matrix(pol, layer_d, layer_n, kp, omega, fromlayer, tolayer, mp)
        return res
    layer_d = list(layer_d)
    layer_n = list(layer_n)
    if fromlayer == 0:
        smat = layersystem_scattering_matrix(pol, layer_d, layer_n, kpar, omega, mp)
        lmat = _mat([[0, 0], [smat[1, 0], 0]], mp)
    elif fromlayer == len(layer_d) - 1:
        smat = layersystem_scattering_matrix(pol, layer_d, layer_n, kpar, omega, mp)
        lmat = _mat([[0, smat[0, 1]], [0, 0]], mp)
    else:
        d_above = [0] + layer_d[fromlayer:]
        n_above = [layer_n[fromlayer]] + layer_n[fromlayer:]
        s_above = layersystem_scattering_matrix(pol, d_above, n_above, kpar, omega, mp)
        d_below = layer_d[:fromlayer] + [0]
        n_below = layer_n[:fromlayer] + [layer_n[fromlayer]]
        s_below = layersystem_scattering_matrix(pol, d_below, n_below, kpar, omega, mp)
        lmat = _mul(_inv(_mat([[1, -s_below[0, 1]], [-s_above[1, 0], 1]], mp), mp),
                    _mat([[0, s_below[0, 1]], [s_above[1, 0], 0]], mp), mp)
    if tolayer > fromlayer:
        tft = layersystem_transfer_matrix(pol, layer_d[fromlayer:tolayer + 1], layer_n[fromlayer:tolayer + 1],
                                          kpar, omega, mp)
        lmat = _mul(_inv(tft, mp), lmat + _mat([[1, 0], [0, 0]], mp), mp)
    elif tolayer < fromlayer:
        tft = layersystem_transfer_matrix(pol, layer_d[tolayer:fromlayer + 1], layer_n[tolayer:fromlayer + 1],
                                          kpar, omega, mp)
        lmat = _mul(tft, lmat + _mat([[0, 0], [0, 1]], mp), mp)
    if mp is not None:
        return np.array([[complex(lmat[0, 0]), complex(lmat[0, 1])], [complex(lmat[1, 0]), complex(lmat[1, 1])]])
    return np.array(lmat, dtype=complex)


def mie_coefficient(tau, l, k_medium, k_particle, radius):
    """Sphere Mie coefficient Q_{tau l} (t_matrix.py:8-38)."""
    jm = fb.spherical_bessel(l, k_medium * radius)
    jp = fb.spherical_bessel(l, k_particle * radius)
    dxjm = fb.dx_xj(l, k_medium * radius)
    dxjp = fb.dx_xj(l, k_particle * radius)
    hm = fb.spherical_hankel(l, k_medium * radius)
    dxhm = fb.dx_xh(l, k_medium * radius)
    if tau == 0:
        return (jm * dxjp - jp * dxjm) / (jp * dxhm - hm * dxjp)
    return ((k_medium ** 2 * jm * dxjp - k_particle ** 2 * jp * dxjm)
            / (k_particle ** 2 * jp * dxhm - k_medium ** 2 * hm * dxjp))


def t_matrix_sphere(k_medium, k_particle, radius, l_max, m_max):
    """Diagonal sphere T-matrix (t_matrix.py:74-93)."""
    b = fb.blocksize(l_max, m_max)
    t = np.zeros((b, b), dtype=complex)
    for n, tau, l, m in fb.index_specs(l_max, m_max):
        t[n, n] = mie_coefficient(tau, l, k_medium, k_particle, radius)
    return t
