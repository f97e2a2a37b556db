"""Drop-in linear system classes (linearsystem/linear_system.py:42-975) whose
coupling operators and solver run on the MI355X through libsmh.

    (1 - T W) b = T a

W is applied by the lookup-interpolated coupling kernels; T is the block
diagonal of the sphere T-matrices; the solve is GMRES on the GPU (FP64, or
FP32 inner iterations under FP64 iterative refinement).
"""
import ctypes
import sys
import time
import warnings

import numpy as np
import scipy.sparse.linalg

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import lookup as look

_INTERP = {'linear': _lib.INTERP_LINEAR, 'catmull-rom': _lib.INTERP_CATMULL_ROM, 'spline': _lib.INTERP_SPLINE}


def _interp_code(interpolator_kind, cubic_interpolant):
    if interpolator_kind == 'linear':
        return _lib.INTERP_LINEAR
    if interpolator_kind != 'cubic':
        raise ValueError('interpolator_kind must be linear or cubic')
    if cubic_interpolant not in ('spline', 'catmull-rom'):
        raise ValueError("cubic_interpolant must be 'spline' or 'catmull-rom'")
    return _INTERP[cubic_interpolant]


class SystemMatrix:
    """Operator on system vectors: particle after particle, blocksize(l_max, m_max) each."""

    def __init__(self, particle_list):
        self.particle_list = particle_list
        self.blocksizes = np.array([flds.blocksize(p.l_max, p.m_max) for p in particle_list], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.blocksizes)])
        n = int(self.offsets[-1])
        self.shape = (n, n)

    def index_block(self, i):
        return range(int(self.offsets[i]), int(self.offsets[i + 1]))

    def index(self, i, tau, l, m):
        p = self.particle_list[i]
        return int(self.offsets[i]) + flds.multi_to_single_index(tau, l, m, p.l_max, p.m_max)


class _GpuCouplingOperator(SystemMatrix):
    """Common part of the GPU lookup operators: the libsmh handle, matvec and timing."""

    _op = None

    def _init_common(self, particle_list):
        SystemMatrix.__init__(self, particle_list)
        self.l_max = max(p.l_max for p in particle_list)
        self.m_max = max(p.m_max for p in particle_list)
        self.blocksize = flds.blocksize(self.l_max, self.m_max)
        self._lmax = np.ascontiguousarray([p.l_max for p in particle_list], dtype=np.int32)
        self._mmax = np.ascontiguousarray([p.m_max for p in particle_list], dtype=np.int32)

    def matvec(self, vector, precision='single'):
        """W x for one vector or a [nrhs, n] block (complex128 in and out)."""
        v = np.ascontiguousarray(vector, dtype=np.complex128)
        nrhs = 1 if v.ndim == 1 else v.shape[0]
        out = np.empty_like(v)
        prec = _lib.FP32 if precision in ('single', 'fp32', _lib.FP32) else _lib.FP64
        _lib.check(_lib.load().smh_op_matvec(self._op, _lib.dptr(v), _lib.dptr(out), nrhs, prec), cu.context(),
                   'smh_op_matvec')
        return out

    def bench_matvec(self, precision='single', iters=10):
        """(kernel ms, matvec ms) averaged over iters device-resident matvecs."""
        k = ctypes.c_double()
        t = ctypes.c_double()
        prec = _lib.FP32 if precision in ('single', 'fp32') else _lib.FP64
        _lib.check(_lib.load().smh_op_bench_matvec(self._op, prec, iters, ctypes.byref(k), ctypes.byref(t)),
                   cu.context(), 'smh_op_bench_matvec')
        return k.value, t.value

    def _make_linear_operator(self):
        self.linear_operator = scipy.sparse.linalg.LinearOperator(shape=self.shape, matvec=self.matvec,
                                                                  dtype=complex)

    def __del__(self):
        if self._op is not None and _lib._LIB is not None:
            _lib.load().smh_op_destroy(self._op)
            self._op = None


class CouplingMatrixRadialLookupCUDA(_GpuCouplingOperator):
    """Radial-lookup coupling matrix on the MI355X; drop-in for smuthi's class of
    the same name (linear_system.py:813-900). Particles must share one height.

    Extra keyword: cubic_interpolant -- 'spline' (default) evaluates the
    not-a-knot spline of smuthi's CPU path (scipy interp1d 'cubic'),
    'catmull-rom' the 4-tap Hermite interpolant of smuthi's CUDA kernel."""

    def __init__(self, vacuum_wavelength, particle_list, layer_system, k_parallel='default', resolution=None,
                 cuda_blocksize=None, interpolator_kind='linear', cubic_interpolant='spline'):
        z = [p.position[2] for p in particle_list]
        if z.count(z[0]) != len(z):
            raise ValueError('radial lookup needs all particles at the same height')
        self._init_common(particle_list)
        if resolution is None:
            resolution = vacuum_wavelength / 100
        self.resolution = resolution
        self.interpolator_kind = interpolator_kind
        interp = _interp_code(interpolator_kind, cubic_interpolant)
        if isinstance(k_parallel, str) and k_parallel == 'default':
            k_parallel = flds.default_Sommerfeld_k_parallel_array
        xy = np.ascontiguousarray([[p.position[0], p.position[1]] for p in particle_list], dtype=np.float64)
        rho_max, rho_min_pair = look.pair_distance_extremes(xy)
        self.radial_distance_array = look.radial_grid(rho_max, resolution)
        i_s = layer_system.layer_number(z[0])
        k_is = complex(layer_system.wavenumber(i_s, vacuum_wavelength))
        dz = z[0] - layer_system.reference_z(i_s)

        lib = _lib.load()
        ctx = cu.context()
        handle = ctypes.c_void_p()
        _lib.check(lib.smh_radial_create(ctx, len(particle_list), _lib.dptr(xy), _lib.iptr(self._lmax),
                                         _lib.iptr(self._mmax), self.l_max, self.m_max,
                                         float(self.radial_distance_array[0]), float(resolution),
                                         len(self.radial_distance_array), interp, ctypes.byref(handle)), ctx,
                   'smh_radial_create')
        self._op = handle
        t0 = time.time()
        coef = np.ascontiguousarray(look.direct_coefficients_radial(self.l_max, self.m_max))
        lm = kp = None
        if not layer_system.is_degenerate():
            kp = np.ascontiguousarray(k_parallel, dtype=np.complex128)
            lm = np.ascontiguousarray(look.layer_coefficients_radial(vacuum_wavelength, layer_system, kp, i_s, dz,
                                                                     self.l_max, self.m_max))
        _lib.check(lib.smh_radial_build_table(self._op, k_is.real, k_is.imag, rho_min_pair / 2, _lib.dptr(coef),
                                              0 if kp is None else len(kp), _lib.dptr(kp), _lib.dptr(lm)), ctx,
                   'smh_radial_build_table')
        self.table_seconds = time.time() - t0
        self._make_linear_operator()

    @property
    def lookup_table(self):
        """The FP64 table [len_rho, B, B] (copied from the device on access)."""
        out = np.empty((len(self.radial_distance_array), self.blocksize, self.blocksize), dtype=np.complex128)
        _lib.check(_lib.load().smh_op_download_table(self._op, _lib.dptr(out), None), cu.context(),
                   'smh_op_download_table')
        return out

    @classmethod
    def from_table(cls, table, radial_distance_array, particle_list, resolution, interpolator_kind='linear',
                   cubic_interpolant='spline'):
        """Operator over a given host table (used to run the kernels on an
        externally computed table, e.g. smuthi's own)."""
        self = cls.__new__(cls)
        self._init_common(particle_list)
        self.resolution = resolution
        self.interpolator_kind = interpolator_kind
        self.radial_distance_array = np.asarray(radial_distance_array)
        interp = _interp_code(interpolator_kind, cubic_interpolant)
        xy = np.ascontiguousarray([[p.position[0], p.position[1]] for p in particle_list], dtype=np.float64)
        lib = _lib.load()
        ctx = cu.context()
        handle = ctypes.c_void_p()
        _lib.check(lib.smh_radial_create(ctx, len(particle_list), _lib.dptr(xy), _lib.iptr(self._lmax),
                                         _lib.iptr(self._mmax), self.l_max, self.m_max,
                                         float(self.radial_distance_array[0]), float(resolution),
                                         len(self.radial_distance_array), interp, ctypes.byref(handle)), ctx,
                   'smh_radial_create')
        self._op = handle
        tab = np.ascontiguousarray(table, dtype=np.complex128)
        _lib.check(lib.smh_op_upload_table(self._op, _lib.dptr(tab), None), ctx, 'smh_op_upload_table')
        self.table_seconds = 0.0
        self._make_linear_operator()
        return self
