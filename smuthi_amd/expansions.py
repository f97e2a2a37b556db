"""Field expansions with smuthi's interface (fields/expansions.py:242-1155):
plane wave expansions whose E/H fields are evaluated by the libsmh PWE
kernels, spherical wave expansion containers, and the layer response of a
PWE (layers.py:141-191).
"""
import ctypes

import numpy as np

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import layers as lay


class FieldExpansion:
    def __init__(self):
        self.validity_conditions = []

    def valid(self, x, y, z):
        ret = np.ones(np.shape(x), dtype=bool)
        for check in self.validity_conditions:
            ret = np.logical_and(ret, check(x, y, z))
        return ret


class PiecewiseFieldExpansion(FieldExpansion):
    """Sum of field expansions, each contributing where it is valid
    (fields/expansions.py:102-195)."""

    def __init__(self):
        FieldExpansion.__init__(self)
        self.expansion_list = []

    def valid(self, x, y, z):
        vld = np.zeros(np.shape(x), dtype=bool)
        for fex in self.expansion_list:
            vld = np.logical_or(vld, fex.valid(x, y, z))
        return np.logical_and(vld, FieldExpansion.valid(self, x, y, z))

    def _sum(self, fn, x, y, z):
        x, y, z = np.array(x, ndmin=1), np.array(y, ndmin=1), np.array(z, ndmin=1)
        out = [np.zeros(x.shape, dtype=complex) for _ in range(3)]
        vld = self.valid(x, y, z)
        for fex in self.expansion_list:
            d = fn(fex, x, y, z)
            for c in range(3):
                out[c][vld] += d[c][vld]
        return tuple(out)

    def electric_field(self, x, y, z, precision='single'):
        return self._sum(lambda f, *r: f.electric_field(*r, precision=precision), x, y, z)

    def magnetic_field(self, x, y, z, vacuum_wavelength, precision='single'):
        return self._sum(lambda f, *r: f.magnetic_field(*r, vacuum_wavelength, precision=precision), x, y, z)


class SphericalWaveExpansion(FieldExpansion):
    """Coefficients a_{tau l m} of a regular or outgoing SWE about reference_point."""

    def __init__(self, k, l_max, m_max=None, kind=None, reference_point=None, lower_z=-np.inf, upper_z=np.inf,
                 inner_r=0, outer_r=np.inf):
        FieldExpansion.__init__(self)
        self.k = k
        self.l_max = l_max
        self.m_max = l_max if m_max is None else m_max
        self.coefficients = np.zeros(flds.blocksize(self.l_max, self.m_max), dtype=complex)
        self.kind = kind
        self.reference_point = reference_point
        self.lower_z, self.upper_z = lower_z, upper_z
        self.inner_r, self.outer_r = inner_r, outer_r

    def coefficients_tlm(self, tau, l, m):
        return self.coefficients[flds.multi_to_single_index(tau, l, m, self.l_max, self.m_max)]

    def valid(self, x, y, z):
        z = np.asarray(z)
        vld = np.logical_and(z >= self.lower_z, z < self.upper_z)
        return np.logical_and(vld, FieldExpansion.valid(self, x, y, z))

    def _field(self, field, x, y, z, omega, precision):
        if self.kind != 'outgoing':
            raise NotImplementedError('the GPU evaluates outgoing spherical wave expansions')
        x, y, z = np.array(x, ndmin=1), np.array(y, ndmin=1), np.array(z, ndmin=1)
        out = [np.zeros(x.shape, dtype=complex) for _ in range(3)]
        vld = self.valid(x, y, z)
        if vld.any():
            pts = np.stack([x[vld], y[vld], z[vld]], axis=1)
            res = swe_field_sum(np.asarray([self.reference_point], dtype=float), np.zeros(1),
                                np.array([[-np.inf, np.inf]]), self.coefficients[None, :], self.l_max, self.m_max,
                                self.k, pts, field, omega, precision)
            for c in range(3):
                out[c][vld] = res[c]
        return tuple(out)

    def electric_field(self, x, y, z, precision='single'):
        """E at points (fields/expansions.py:348-381), zero where not valid."""
        return self._field(0, x, y, z, 1.0, precision)

    def magnetic_field(self, x, y, z, vacuum_wavelength, precision='single'):
        """H at points (fields/expansions.py:383-424), zero where not valid."""
        return self._field(1, x, y, z, flds.angular_frequency(vacuum_wavelength), precision)


def swe_field_sum(positions, radii, zlim, coefficients, l_max, m_max, k, points, field=0, omega=1.0,
                  precision='single'):
    """Sum of outgoing SWEs at points on the GPU (libsmh smh_swe_field):
    positions [n][3], radii [n], zlim [n][2], coefficients [n][blocksize(l_max,
    m_max)]; particle p counts at points with zlim[p, 0] <= z < zlim[p, 1]
    outside its sphere. field 0: E, 1: H. Returns (Fx, Fy, Fz)."""
    pos = np.ascontiguousarray(positions, dtype=np.float64)
    rad = np.ascontiguousarray(radii, dtype=np.float64)
    zl = np.ascontiguousarray(zlim, dtype=np.float64)
    cf = np.ascontiguousarray(coefficients, dtype=np.complex128)
    pts = np.ascontiguousarray(points, dtype=np.float64)
    n_pts = pts.shape[0]
    out = np.empty((3, n_pts), dtype=np.complex128)
    ms = ctypes.c_double()
    prec = _lib.FP32 if precision in ('single', 'fp32') else _lib.FP64
    kk = complex(k)
    _lib.check(_lib.load().smh_swe_field(cu.context(), len(pos), _lib.dptr(pos), _lib.dptr(rad), _lib.dptr(zl),
                                         int(l_max), int(m_max), _lib.dptr(cf), kk.real, kk.imag, n_pts,
                                         _lib.dptr(pts), int(field), float(omega), prec, _lib.dptr(out),
                                         ctypes.byref(ms)), cu.context(), 'smh_swe_field')
    swe_field_sum.last_kernel_ms = ms.value
    return out[0], out[1], out[2]


class SphericalWaveSum(FieldExpansion):
    """The outgoing scattered-field SWEs of the particles of one layer as one
    expansion: the direct field that PiecewiseFieldExpansion adds particle by
    particle (postprocessing/scattered_field.py:72-75), summed on the GPU.
    Each SWE counts in [lower_z, upper_z) outside its own sphere."""

    def __init__(self, k, particles, lower_z=-np.inf, upper_z=np.inf):
        FieldExpansion.__init__(self)
        self.k = k
        self.lower_z, self.upper_z = lower_z, upper_z
        # a particle's scattered field: a SphericalWaveExpansion, or its coefficient array
        orders = [(getattr(p.scattered_field, 'l_max', p.l_max), getattr(p.scattered_field, 'm_max', p.m_max))
                  for p in particles]
        self.l_max = max(o[0] for o in orders)
        self.m_max = max(o[1] for o in orders)
        B = flds.blocksize(self.l_max, self.m_max)
        self.positions = np.array([p.position for p in particles], dtype=float)
        self.radii = np.array([p.radius for p in particles], dtype=float)
        self.coefficients = np.zeros((len(particles), B), dtype=complex)
        for i, (p, (lp, mp)) in enumerate(zip(particles, orders)):
            cp = np.asarray(getattr(p.scattered_field, 'coefficients', p.scattered_field), dtype=complex)
            for tau in range(2):
                for l in range(1, lp + 1):
                    for m in range(-min(l, mp), min(l, mp) + 1):
                        self.coefficients[i, flds.multi_to_single_index(tau, l, m, self.l_max, self.m_max)] = \
                            cp[flds.multi_to_single_index(tau, l, m, lp, mp)]
        self._particles = particles

    def valid(self, x, y, z):
        z = np.asarray(z)
        vld = np.logical_and(z >= self.lower_z, z < self.upper_z)
        if len(self._particles) == 1:
            vld = np.logical_and(vld, self._particles[0].is_outside(x, y, z))
        return np.logical_and(vld, FieldExpansion.valid(self, x, y, z))

    def _field(self, field, x, y, z, omega, precision):
        x, y, z = np.array(x, ndmin=1), np.array(y, ndmin=1), np.array(z, ndmin=1)
        pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
        zlim = np.tile([self.lower_z, self.upper_z], (len(self.radii), 1))
        res = swe_field_sum(self.positions, self.radii, zlim, self.coefficients, self.l_max, self.m_max, self.k,
                            pts, field, omega, precision)
        return tuple(r.reshape(x.shape) for r in res)

    def electric_field(self, x, y, z, precision='single'):
        return self._field(0, x, y, z, 1.0, precision)

    def magnetic_field(self, x, y, z, vacuum_wavelength, precision='single'):
        return self._field(1, x, y, z, flds.angular_frequency(vacuum_wavelength), precision)


class PlaneWaveExpansion(FieldExpansion):
    """g_j(kappa, alpha) on a (k_parallel, azimuthal_angles) grid; a single
    entry in each means a discrete plane wave."""

    def __init__(self, k, k_parallel, azimuthal_angles, kind=None, reference_point=None, lower_z=-np.inf,
                 upper_z=np.inf):
        FieldExpansion.__init__(self)
        self.k = k
        self.k_parallel = np.array(k_parallel, ndmin=1)
        self.azimuthal_angles = np.array(azimuthal_angles, ndmin=1)
        self.kind = kind
        self.reference_point = reference_point
        self.lower_z, self.upper_z = lower_z, upper_z
        self.coefficients = np.zeros((2, len(self.k_parallel), len(self.azimuthal_angles)), dtype=complex)

    def k_z(self):
        kz = flds.k_z(k_parallel=self.k_parallel, k=self.k)
        if self.kind == 'upgoing':
            return kz
        if self.kind == 'downgoing':
            return -kz
        raise ValueError('pwe kind undefined')

    def k_parallel_grid(self):
        return np.meshgrid(self.k_parallel, self.azimuthal_angles, indexing='ij')[0]

    def azimuthal_angle_grid(self):
        return np.meshgrid(self.k_parallel, self.azimuthal_angles, indexing='ij')[1]

    def k_z_grid(self):
        return np.broadcast_to(self.k_z()[:, None], (len(self.k_parallel), len(self.azimuthal_angles)))

    def valid(self, x, y, z):
        z = np.asarray(z)
        if self.upper_z == float('inf'):
            vld = z >= self.lower_z
        else:
            vld = np.logical_and(z >= self.lower_z, z < self.upper_z)
        return np.logical_and(vld, FieldExpansion.valid(self, x, y, z))

    def compatible(self, other):
        return (isinstance(other, PlaneWaveExpansion) and np.isclose(self.k, other.k)
                and len(self.k_parallel) == len(other.k_parallel) and np.allclose(self.k_parallel, other.k_parallel)
                and len(self.azimuthal_angles) == len(other.azimuthal_angles)
                and np.allclose(self.azimuthal_angles, other.azimuthal_angles) and self.kind == other.kind
                and np.array_equal(self.reference_point, other.reference_point))

    def __add__(self, other):
        if not self.compatible(other):
            raise ValueError('Plane wave expansion are inconsistent.')
        out = PlaneWaveExpansion(self.k, self.k_parallel, self.azimuthal_angles, self.kind, self.reference_point,
                                 max(self.lower_z, other.lower_z), min(self.upper_z, other.upper_z))
        out.coefficients = self.coefficients + other.coefficients
        return out

    def _field(self, kind, x, y, z, omega, precision):
        x, y, z = (np.asarray(v, dtype=float) for v in (x, y, z))
        out = [np.zeros(x.shape, dtype=complex) for _ in range(3)]
        vld = self.valid(x, y, z)
        if not vld.any():
            return tuple(out)
        pts = np.ascontiguousarray(np.stack([x[vld] - self.reference_point[0], y[vld] - self.reference_point[1],
                                             z[vld] - self.reference_point[2]], axis=-1))
        kp = np.ascontiguousarray(self.k_parallel, dtype=np.complex128)
        kz = np.ascontiguousarray(self.k_z(), dtype=np.complex128)
        al = np.ascontiguousarray(self.azimuthal_angles, dtype=np.float64)
        g = np.ascontiguousarray(self.coefficients, dtype=np.complex128)
        res = np.zeros((3, len(pts)), dtype=np.complex128)
        ms = ctypes.c_double()
        prec = _lib.FP64 if precision in ('double', 'double precision') else _lib.FP32
        k = complex(self.k)
        _lib.check(_lib.load().smh_pwe_field(cu.context(), kind, len(kp), _lib.dptr(kp), _lib.dptr(kz), len(al),
                                              _lib.dptr(al), _lib.dptr(g), k.real, k.imag, len(pts), _lib.dptr(pts),
                                              _lib.dptr(res), prec, ctypes.byref(ms)), cu.context(), 'smh_pwe_field')
        self.last_kernel_ms = ms.value
        for c in range(3):
            out[c][vld] = res[c] / omega
        return tuple(out)

    def electric_field(self, x, y, z, max_chunksize=50, cpu_precision='single precision', precision='single'):
        """(E_x, E_y, E_z) at the points; zero where the expansion is not valid."""
        return self._field(0, x, y, z, 1.0, precision)

    def magnetic_field(self, x, y, z, vacuum_wavelength, max_chunksize=50, cpu_precision='single precision',
                       precision='single'):
        """(H_x, H_y, H_z) at the points; zero where the expansion is not valid."""
        return self._field(1, x, y, z, flds.angular_frequency(vacuum_wavelength), precision)


def layer_response(layer_system, pwe, from_layer, to_layer):
    """(upgoing, downgoing) PWEs of the layer system response to pwe (or to a (up, down) tuple)."""
    if isinstance(pwe, (tuple, list)):
        u0, d0 = layer_response(layer_system, pwe[0], from_layer, to_layer)
        u1, d1 = layer_response(layer_system, pwe[1], from_layer, to_layer)
        return u0 + u1, d0 + d1
    omega = pwe.k / layer_system.refractive_indices[from_layer]
    k_to = omega * layer_system.refractive_indices[to_layer]
    ref = [0, 0, layer_system.reference_z(to_layer)]
    loz, upz = layer_system.lower_zlimit(to_layer), layer_system.upper_zlimit(to_layer)
    up = PlaneWaveExpansion(k_to, pwe.k_parallel, pwe.azimuthal_angles, 'upgoing', ref, loz, upz)
    dn = PlaneWaveExpansion(k_to, pwe.k_parallel, pwe.azimuthal_angles, 'downgoing', ref, loz, upz)
    col = 0 if pwe.kind == 'upgoing' else 1
    for pol in range(2):
        L = lay.layersystem_response_matrix(pol, layer_system.thicknesses, layer_system.refractive_indices,
                                            pwe.k_parallel, omega, from_layer, to_layer)
        up.coefficients[pol] = L[0, col][:, None] * pwe.coefficients[pol]
        dn.coefficients[pol] = L[1, col][:, None] * pwe.coefficients[pol]
    return up, dn
