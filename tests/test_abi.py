"""The C-ABI library builds, loads and exports every symbol of include/smh.h."""
import os
import re

import numpy as np
import pytest
import scipy.special

from smuthi_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, 'include', 'smh.h')).read()
    return sorted(set(re.findall(r'\b(smh_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, name


def test_host_bessel_matches_scipy():
    rng = np.random.default_rng(5)
    # arguments as they occur on Sommerfeld contours: |z| up to ~2000, small imaginary parts,
    # negative real parts for the rho < 0 padding rows, plus the branch regions of the algorithm
    mag = np.concatenate([rng.uniform(0, 4, 300), rng.uniform(4, 25, 300), rng.uniform(25, 2500, 300),
                          [1e-12, 3.999, 4.001, 24.99, 25.01]])
    ang = rng.uniform(-0.12, 0.05, mag.size)
    z = mag * np.exp(1j * ang)
    z[::7] *= -1
    nmax = 8
    out = np.zeros((z.size, nmax + 1), complex)
    _lib.check(_lib.load().smh_host_bessel_j(nmax, z.size, _lib.dptr(np.ascontiguousarray(z)), _lib.dptr(out)))
    for n in range(nmax + 1):
        ref = scipy.special.jv(n, z)
        scale = np.maximum(np.abs(ref), np.abs(scipy.special.hankel1(n, z)) * 1e-3)
        err = np.abs(out[:, n] - ref) / scale
        assert err.max() < 1e-12, (n, z[np.argmax(err)], err.max())


def test_host_spherical_hankel_matches_scipy():
    rng = np.random.default_rng(6)
    z = rng.uniform(0.5, 3000, 500) * np.exp(1j * rng.uniform(-0.01, 0.01, 500))
    z = np.concatenate([z, [0.3, 1.0, 2.5]]).astype(complex)
    nmax = 8
    out = np.zeros((z.size, nmax + 1), complex)
    _lib.check(_lib.load().smh_host_spherical_hankel(nmax, z.size, _lib.dptr(np.ascontiguousarray(z)),
                                                     _lib.dptr(out)))
    for n in range(nmax + 1):
        ref = np.sqrt(np.pi / (2 * z)) * scipy.special.hankel1(n + 0.5, z)
        err = np.abs(out[:, n] - ref) / np.abs(ref)
        assert err.max() < 1e-12, (n, err.max())
