"""Angular functions and SVWF/PVWF coefficients on the host (setup only):
normalized associated Legendre P, pi, tau of complex argument
(utility/math.py:38-104), Wigner-3j (utility/math.py:14-35), a5/b5 translation
coefficients (fields/transformations.py:413-443) and the transformation
coefficients B, B-dagger (fields/transformations.py:18-74), all vectorized
over the argument arrays.
"""
import math
from functools import lru_cache

import numpy as np

from . import fields as flds


def legendre_normalized(ct, st, lmax):
    """P_l^m, pi_l^m, tau_l^m arrays indexed [l, m, ...] for complex ct, st."""
    ct = np.asarray(ct, dtype=complex)
    st = np.asarray(st, dtype=complex)
    shp = (lmax + 1, lmax + 1) + ct.shape
    P = np.zeros(shp, complex)
    PI = np.zeros(shp, complex)
    TAU = np.zeros(shp, complex)
    dP0 = np.zeros((lmax + 1,) + ct.shape, complex)  # d/dct of P_l^0
    P[0, 0] = np.sqrt(0.5)
    TAU[0, 0] = 0.0
    if lmax >= 1:
        P[1, 0] = np.sqrt(1.5) * ct
        dP0[1] = np.sqrt(1.5)
        TAU[1, 0] = -st * dP0[1]
    for l in range(1, lmax):
        a = np.sqrt((2 * l + 1) * (2 * l + 3)) / (l + 1)
        b = l / (l + 1) * np.sqrt((2 * l + 3) / (2 * l - 1))
        P[l + 1, 0] = a * ct * P[l, 0] - b * P[l - 1, 0]
        c = np.sqrt((2 * l + 3) / (2 * l + 1))
        dP0[l + 1] = c * ((l + 1) * P[l, 0] + ct * dP0[l])
        TAU[l + 1, 0] = -st * dP0[l + 1]
    for m in range(1, lmax + 1):
        # sqrt((2m+1)!! / (2 (2m)!!)) as a product of (t+1)/t over even t
        pref = math.sqrt(math.prod((t + 1) / t for t in range(2, 2 * m + 1, 2)) / 2)
        P[m, m] = pref * st ** m
        PI[m, m] = pref * st ** (m - 1)
        TAU[m, m] = m * ct * PI[m, m]
        for l in range(m, lmax):
            den = (l + 1 - m) * (l + 1 + m)
            a = np.sqrt((2 * l + 1) * (2 * l + 3) / den)
            b = np.sqrt((2 * l + 3) * (l - m) * (l + m) / ((2 * l - 1) * den))
            P[l + 1, m] = a * ct * P[l, m] - b * P[l - 1, m]
            PI[l + 1, m] = a * ct * PI[l, m] - b * PI[l - 1, m]
            TAU[l + 1, m] = ((l + 1) * ct * PI[l + 1, m]
                             - (l + 1 + m) * np.sqrt((2 * l + 3) * (l + 1 - m) / ((2 * l + 1) * (l + 1 + m)))
                             * PI[l, m])
    return P, PI, TAU


@lru_cache(maxsize=None)
def wigner3j(j1, j2, j3, m1, m2, m3):
    """Wigner-3j for integer arguments: Racah sum evaluated with log-factorials."""
    j1, j2, j3, m1, m2, m3 = (int(v) for v in (j1, j2, j3, m1, m2, m3))
    if m1 + m2 + m3 != 0 or j3 < abs(j1 - j2) or j3 > j1 + j2 or abs(m1) > j1 or abs(m2) > j2 or abs(m3) > j3:
        return 0.0
    lf = math.lgamma
    log_tri = 0.5 * (lf(j1 + j2 - j3 + 1) + lf(j1 - j2 + j3 + 1) + lf(-j1 + j2 + j3 + 1) - lf(j1 + j2 + j3 + 2))
    log_pre = 0.5 * (lf(j1 + m1 + 1) + lf(j1 - m1 + 1) + lf(j2 + m2 + 1) + lf(j2 - m2 + 1)
                     + lf(j3 + m3 + 1) + lf(j3 - m3 + 1))
    total = 0.0
    for k in range(max(0, j2 - j3 - m1, j1 - j3 + m2), min(j1 + j2 - j3, j1 - m1, j2 + m2) + 1):
        lden = (lf(k + 1) + lf(j1 + j2 - j3 - k + 1) + lf(j1 - m1 - k + 1) + lf(j2 + m2 - k + 1)
                + lf(j3 - j2 + m1 + k + 1) + lf(j3 - j1 - m2 + k + 1))
        total += (-1) ** k * math.exp(log_tri + log_pre - lden)
    return (-1) ** (j1 - j2 - m3) * total


@lru_cache(maxsize=None)
def ab5_coefficients(l1, m1, l2, m2, p):
    """a5, b5 of the SVWF translation operator."""
    l1, m1, l2, m2, p = int(l1), int(m1), int(l2), int(m2), int(p)
    jfac = 1j ** (abs(m1 - m2) - abs(m1) - abs(m2) + l2 - l1 + p) * (-1) ** (m1 - m2)
    fac1 = math.sqrt((2 * l1 + 1) * (2 * l2 + 1) / (2 * l1 * (l1 + 1) * l2 * (l2 + 1)))
    fac2a = (l1 * (l1 + 1) + l2 * (l2 + 1) - p * (p + 1)) * math.sqrt(2 * p + 1)
    fac2b = math.sqrt(max(0, (l1 + l2 + 1 + p) * (l1 + l2 + 1 - p) * (p + l1 - l2) * (p - l1 + l2) * (2 * p + 1)))
    w1 = wigner3j(l1, l2, p, m1, -m2, -(m1 - m2))
    a = jfac * fac1 * fac2a * w1 * wigner3j(l1, l2, p, 0, 0, 0)
    b = jfac * fac1 * fac2b * w1 * (wigner3j(l1, l2, p - 1, 0, 0, 0) if p >= 1 else 0.0)
    return a, b


def transformation_arrays(l_max, m_max, ct, st, dagger):
    """B[pol, n, ...] (or B-dagger) for all multipole indices n of (l_max, m_max)."""
    _, PI, TAU = legendre_normalized(ct, st, l_max)
    tau, ll, mm = flds.index_arrays(l_max, m_max)
    out = np.zeros((2, len(tau)) + np.shape(ct), complex)
    sgn = -1j if dagger else 1j
    for n in range(len(tau)):
        t, l, m = tau[n], ll[n], mm[n]
        base = -1 / sgn ** (l + 1) / math.sqrt(2 * l * (l + 1))
        for pol in range(2):
            fun = TAU[l, abs(m)] if t == pol else m * PI[l, abs(m)]
            out[pol, n] = base * (sgn if pol == 0 else 1) * fun
    return out
