"""Planar layer system (layers.py:17-380): geometry and the layer response
L[pol][+-][+-](kappa), evaluated for whole kappa arrays at once with batched
2x2 scattering/transfer matrices (host setup for tables, RHS and PWEs).
"""
import numpy as np

from . import fields as flds


class LayerSystem:
    """Stack of planar layers; first and last are half spaces (thickness 0)."""

    def __init__(self, thicknesses=None, refractive_indices=None):
        thicknesses = [0, 0] if thicknesses is None else list(thicknesses)
        refractive_indices = [1, 1] if refractive_indices is None else list(refractive_indices)
        if len(thicknesses) != len(refractive_indices):
            raise ValueError('each layer needs one refractive index')
        self.thicknesses = list(thicknesses)
        self.thicknesses[0] = 0
        self.thicknesses[-1] = 0
        self.refractive_indices = list(refractive_indices)

    def is_degenerate(self):
        return self.thicknesses == [0, 0] and self.refractive_indices[0] == self.refractive_indices[-1]

    def number_of_layers(self):
        return len(self.thicknesses)

    def lower_zlimit(self, i):
        return -np.inf if i == 0 else float(sum(self.thicknesses[1:i]))

    def upper_zlimit(self, i):
        return np.inf if i == self.number_of_layers() - 1 else float(sum(self.thicknesses[1:i + 1]))

    def reference_z(self, i):
        return self.upper_zlimit(i) if i == 0 else self.lower_zlimit(i)

    def layer_number(self, z):
        """Layer containing height z; a point on an interface belongs to the upper layer."""
        bounds = np.cumsum([0] + list(self.thicknesses[1:-1]))
        return int(np.searchsorted(bounds, z, side='right'))

    def wavenumber(self, layer_number, vacuum_wavelength):
        return self.refractive_indices[layer_number] * flds.angular_frequency(vacuum_wavelength)

    def response(self, pwe, from_layer, to_layer):
        from .expansions import layer_response
        return layer_response(self, pwe, from_layer, to_layer)


def _mm(a, b):
    """Batched 2x2 product over the trailing axis."""
    return np.einsum('ij...,jk...->ik...', a, b)


def _inv(a):
    det = a[0, 0] * a[1, 1] - a[0, 1] * a[1, 0]
    return np.array([[a[1, 1], -a[0, 1]], [-a[1, 0], a[0, 0]]]) / det


def _layer_kz(layer_n, kpar, omega):
    return [flds.k_z(k_parallel=kpar, k=omega * n) for n in layer_n]


def _interface(pol, kz1, kz2, n1, n2):
    if pol == 0:
        r = (kz1 - kz2) / (kz1 + kz2)
        t = 2 * kz1 / (kz1 + kz2)
    else:
        r = (n2 ** 2 * kz1 - n1 ** 2 * kz2) / (n2 ** 2 * kz1 + n1 ** 2 * kz2)
        t = 2 * n1 * n2 * kz1 / (n2 ** 2 * kz1 + n1 ** 2 * kz2)
    return np.array([[1 / t, r / t], [r / t, 1 / t]])


def _propagation(kz, d):
    z = np.zeros_like(kz)
    return np.array([[np.exp(-1j * kz * d), z], [z, np.exp(1j * kz * d)]])


def transfer_matrix(pol, layer_d, layer_n, kpar, omega):
    kzs = _layer_kz(layer_n, kpar, omega)
    one = np.ones_like(kzs[0])
    zero = np.zeros_like(kzs[0])
    t = np.array([[one, zero], [zero, one]])
    for i in range(len(layer_d) - 1):
        t = _mm(t, _mm(_propagation(kzs[i], layer_d[i]), _interface(pol, kzs[i], kzs[i + 1], layer_n[i],
                                                                    layer_n[i + 1])))
    return t


def scattering_matrix(pol, layer_d, layer_n, kpar, omega):
    kzs = _layer_kz(layer_n, kpar, omega)
    one = np.ones_like(kzs[0])
    zero = np.zeros_like(kzs[0])
    s = np.array([[one, zero], [zero, one]])
    for i in range(len(layer_d) - 1):
        t = _mm(_propagation(kzs[i], layer_d[i]), _interface(pol, kzs[i], kzs[i + 1], layer_n[i], layer_n[i + 1]))
        den = t[0, 0] - s[0, 1] * t[1, 0]
        s11 = s[0, 0] / den
        s12 = (s[0, 1] * t[1, 1] - t[0, 1]) / den
        s21 = s[1, 1] * t[1, 0] * s11 + s[1, 0]
        s22 = s[1, 1] * t[1, 0] * s12 + s[1, 1] * t[1, 1]
        s = np.array([[s11, s12], [s21, s22]])
    return s


def layersystem_response_matrix(pol, layer_d, layer_n, kpar, omega, fromlayer, tolayer):
    """Response matrix L[+-, +-, ...] for a scalar or an array of kappa."""
    layer_d, layer_n = list(layer_d), list(layer_n)
    kpar = np.asarray(kpar, dtype=complex)
    zero = np.zeros_like(kpar)
    one = np.ones_like(kpar)
    if fromlayer == 0:
        s = scattering_matrix(pol, layer_d, layer_n, kpar, omega)
        lmat = np.array([[zero, zero], [s[1, 0], zero]])
    elif fromlayer == len(layer_d) - 1:
        s = scattering_matrix(pol, layer_d, layer_n, kpar, omega)
        lmat = np.array([[zero, s[0, 1]], [zero, zero]])
    else:
        sa = scattering_matrix(pol, [0] + layer_d[fromlayer:], [layer_n[fromlayer]] + layer_n[fromlayer:], kpar,
                               omega)
        sb = scattering_matrix(pol, layer_d[:fromlayer] + [0], layer_n[:fromlayer] + [layer_n[fromlayer]], kpar,
                               omega)
        lmat = _mm(_inv(np.array([[one, -sb[0, 1]], [-sa[1, 0], one]])), np.array([[zero, sb[0, 1]],
                                                                                 [sa[1, 0], zero]]))
    if tolayer > fromlayer:
        t = transfer_matrix(pol, layer_d[fromlayer:tolayer + 1], layer_n[fromlayer:tolayer + 1], kpar, omega)
        lmat = _mm(_inv(t), lmat + np.array([[one, zero], [zero, zero]]))
    elif tolayer < fromlayer:
        t = transfer_matrix(pol, layer_d[tolayer:fromlayer + 1], layer_n[tolayer:fromlayer + 1], kpar, omega)
        lmat = _mm(t, lmat + np.array([[zero, zero], [zero, one]]))
    return lmat
