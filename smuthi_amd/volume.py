"""Volume-lookup coupling matrix on the MI355X: drop-in for smuthi's
CouplingMatrixVolumeLookupCUDA (linearsystem/linear_system.py:688-783), for
particles at different heights inside one layer. Both tables (over
(rho, z1+z2) and (rho, z1-z2), prepare_lookup.py:16-250) are built on the GPU
and interpolated by the binned MFMA matvec of libsmh: bilinearly, or bicubic
as a tensor not-a-knot spline (smuthi's CPU RectBivariateSpline(kx=ky=3),
linear_system.py:652-663) or tensor Catmull-Rom (smuthi's CUDA kernel,
linear_system_cuda.py:149-256)."""
import ctypes
import time

import numpy as np

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import lookup as look
from .linear_system import _GpuCouplingOperator


def _interp_code(interpolator_kind, cubic_interpolant):
    if interpolator_kind == 'linear':
        return _lib.INTERP_LINEAR
    if interpolator_kind != 'cubic':
        raise ValueError("interpolator_kind must be 'linear' or 'cubic'")
    if cubic_interpolant == 'spline':
        return _lib.INTERP_SPLINE
    if cubic_interpolant == 'catmull-rom':
        return _lib.INTERP_CATMULL_ROM
    raise ValueError("cubic_interpolant must be 'spline' or 'catmull-rom'")


class CouplingMatrixVolumeLookupCUDA(_GpuCouplingOperator):
    """Attributes as smuthi's class: lookup_table_plus / lookup_table_minus
    [len_rho, len_z, B, B] (complex64 values), rho_array, sum_z_array,
    diff_z_array, linear_operator. interpolator_kind 'linear' (bilinear, the
    reference's GPU 'linear' and CPU kx=ky=1 spline) or 'cubic'; extra keyword
    cubic_interpolant: 'spline' (default, tensor not-a-knot spline = smuthi
    CPU RectBivariateSpline(kx=ky=3)) or 'catmull-rom' (smuthi's CUDA
    kernel)."""

    def __init__(self, vacuum_wavelength, particle_list, layer_system, k_parallel='default', resolution=None,
                 cuda_blocksize=None, interpolator_kind='linear', cubic_interpolant='spline'):
        self._interp = _interp_code(interpolator_kind, cubic_interpolant)
        self.cubic_interpolant = cubic_interpolant
        layers = {layer_system.layer_number(p.position[2]) for p in particle_list}
        if len(layers) != 1:
            raise ValueError('volume lookup needs all particles in one layer')
        self._init_common(particle_list)
        if resolution is None:
            resolution = vacuum_wavelength / 100
        self.resolution = resolution
        self.interpolator_kind = interpolator_kind
        if isinstance(k_parallel, str) and k_parallel == 'default':
            k_parallel = flds.default_Sommerfeld_k_parallel_array
        xyz = np.ascontiguousarray([p.position for p in particle_list], dtype=np.float64)
        self.rho_array, self.sum_z_array, self.diff_z_array, rho_min_pair = look.volume_grids(xyz, resolution)
        self._create(xyz)
        i_s = layers.pop()
        k_is = complex(layer_system.wavenumber(i_s, vacuum_wavelength))
        t0 = time.time()
        coef = np.ascontiguousarray(look.direct_coefficients_volume(self.l_max, self.m_max))
        kp = kz = lc = None
        n_k = 0
        if not layer_system.is_degenerate():
            kp = np.ascontiguousarray(k_parallel, dtype=np.complex128)
            lc, kz = look.layer_coefficients_volume(vacuum_wavelength, layer_system, kp, i_s, self.l_max, self.m_max)
            lc, kz = np.ascontiguousarray(lc), np.ascontiguousarray(kz, dtype=np.complex128)
            n_k = len(kp)
        _lib.check(_lib.load().smh_volume_build_table(self._op, k_is.real, k_is.imag, rho_min_pair / 2,
                                                      float(layer_system.reference_z(i_s)), _lib.dptr(coef), n_k,
                                                      _lib.dptr(kp), _lib.dptr(kz), _lib.dptr(lc)), cu.context(),
                   'smh_volume_build_table')
        self.table_seconds = time.time() - t0
        self._make_linear_operator()

    def _create(self, xyz):
        handle = ctypes.c_void_p()
        h = float(self.resolution)
        _lib.check(_lib.load().smh_volume_create(cu.context(), len(xyz), _lib.dptr(xyz), _lib.iptr(self._lmax),
                                                 _lib.iptr(self._mmax), self.l_max, self.m_max,
                                                 float(self.rho_array[0]), float(self.sum_z_array[0]),
                                                 float(self.diff_z_array[0]), h, len(self.rho_array),
                                                 len(self.sum_z_array), self._interp, ctypes.byref(handle)),
                   cu.context(), 'smh_volume_create')
        self._op = handle

    def _tables(self):
        shape = (len(self.rho_array), len(self.sum_z_array), self.blocksize, self.blocksize)
        plus = np.empty(shape, dtype=np.complex128)
        minus = np.empty(shape, dtype=np.complex128)
        _lib.check(_lib.load().smh_op_download_table(self._op, _lib.dptr(plus), _lib.dptr(minus)), cu.context(),
                   'smh_op_download_table')
        return plus, minus

    @property
    def lookup_table_plus(self):
        return self._tables()[0]

    @property
    def lookup_table_minus(self):
        return self._tables()[1]

    @classmethod
    def from_tables(cls, table_plus, table_minus, rho_array, sum_z_array, diff_z_array, particle_list, resolution,
                    interpolator_kind='linear', cubic_interpolant='spline'):
        """Operator over given host tables (e.g. smuthi's own)."""
        self = cls.__new__(cls)
        self._interp = _interp_code(interpolator_kind, cubic_interpolant)
        self.cubic_interpolant = cubic_interpolant
        self._init_common(particle_list)
        self.resolution = resolution
        self.interpolator_kind = interpolator_kind
        self.rho_array, self.sum_z_array, self.diff_z_array = (np.asarray(a) for a in (rho_array, sum_z_array,
                                                                                    diff_z_array))
        xyz = np.ascontiguousarray([p.position for p in particle_list], dtype=np.float64)
        self._create(xyz)
        tp = np.ascontiguousarray(table_plus, dtype=np.complex128)
        tm = np.ascontiguousarray(table_minus, dtype=np.complex128)
        _lib.check(_lib.load().smh_op_upload_table(self._op, _lib.dptr(tp), _lib.dptr(tm)), cu.context(),
                   'smh_op_upload_table')
        self.table_seconds = 0.0
        self._make_linear_operator()
        return self
