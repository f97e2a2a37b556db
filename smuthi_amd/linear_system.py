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


def _precision_code(precision):
    if precision in ('single', 'fp32', _lib.FP32):
        return _lib.FP32
    if precision in ('half', 'fp16', _lib.FP16):
        return _lib.FP16
    return _lib.FP64


def _device_or_host(a):
    """(array, pointer) of a complex128 vector or block: a cuda.DeviceArray as it is, anything else as a
    contiguous numpy array."""
    if isinstance(a, cu.DeviceArray):
        return a, ctypes.cast(a.ptr, _lib.DPTR)
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a, _lib.dptr(a)


def _solve_buffers(rhs, n):
    """Right-hand sides and a solution buffer of the same kind (DeviceArray or numpy array)."""
    b, b_ptr = _device_or_host(rhs)
    if b.shape[-1] != n or b.ndim not in (1, 2):
        raise ValueError('rhs must be [n] or [nrhs, n] with n = %d' % n)
    if isinstance(b, cu.DeviceArray):
        x = cu.DeviceArray.empty(b.shape)
        return b, b_ptr, x, ctypes.cast(x.ptr, _lib.DPTR)
    x = np.empty_like(b)
    return b, b_ptr, x, _lib.dptr(x)


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
        """W x for one vector or a [nrhs, n] block (complex128 in and out).
        precision: 'single' (FP32), 'double' (FP64) or 'half' (FP16 operands,
        FP32 accumulation; radial and volume operators, a few 1e-4 relative error)."""
        v = np.ascontiguousarray(vector, dtype=np.complex128)
        nrhs = 1 if v.ndim == 1 else v.shape[0]
        out = np.empty_like(v)
        prec = _precision_code(precision)
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

    def bench_matvec_batch(self, precision='single', iters=10):
        """(kernel ms, matvec ms) of one batched launch (8 right-hand sides) on device-resident data."""
        k = ctypes.c_double()
        t = ctypes.c_double()
        prec = _precision_code(precision)
        _lib.check(_lib.load().smh_op_bench_matvec_batch(self._op, prec, iters, ctypes.byref(k), ctypes.byref(t)),
                   cu.context(), 'smh_op_bench_matvec_batch')
        return k.value, t.value

    # launch shapes the autotuner tries: (work items per CU, launch order), the order 0 (chunk order),
    # 1 (XCD-dealt by table row) or n >= 2 (n bands of table rows, launched band after band)
    AUTOTUNE_CANDIDATES = (tuple((n, x) for n in (4, 8, 16, 32, 64) for x in (1, 0))
                           + tuple((n, b) for n in (16, 64) for b in (32, 128, 512)))

    def autotune(self, batch=True, precision='half', iters=3, candidates=None):
        """Time the matvec of one shape (batch: the 8-RHS kernel in 'half' / 'single' / 'double'; else
        the single-vector kernel in 'single' / 'double'; kernels and the reduction of their partial sums)
        over launch shapes (work items per CU, launch order: 0 chunk order, 1 XCD-dealt by table row,
        n >= 2 row bands) and keep the fastest (smh_op_autotune). Returns {'best': (items_per_cu, order),
        'matvec_ms': {(items_per_cu, order): ms per matvec}}."""
        cand = list(self.AUTOTUNE_CANDIDATES if candidates is None else candidates)
        per_cu = np.ascontiguousarray([c[0] for c in cand], dtype=np.int32)
        xcd = np.ascontiguousarray([int(c[1]) for c in cand], dtype=np.int32)
        ms = np.zeros(len(cand))
        best = ctypes.c_int(-1)
        _lib.check(_lib.load().smh_op_autotune(self._op, int(bool(batch)), _precision_code(precision), int(iters),
                                               len(cand), _lib.iptr(per_cu), _lib.iptr(xcd), _lib.dptr(ms),
                                               ctypes.byref(best)), cu.context(), 'smh_op_autotune')
        return dict(best=tuple(cand[best.value]), matvec_ms={tuple(c): float(m) for c, m in zip(cand, ms)})

    def half_interpolation(self):
        """(Horner terms, measured linear-interpolation error / bin scale) of the FP16 operator of the
        mixed16 inner solves: 2 terms = linear between the table nodes."""
        terms = ctypes.c_int()
        ratio = ctypes.c_double()
        _lib.check(_lib.load().smh_op_half_interp(self._op, ctypes.byref(terms), ctypes.byref(ratio)), cu.context(),
                   'smh_op_half_interp')
        return terms.value, ratio.value

    def solve_master(self, t_diagonal, rhs, tol, precision='mixed', restart=20, maxiter=None, with_stats=False):
        """Solve (1 - diag(t) W) x = rhs on the GPU (rhs [n] or [nrhs, n]).

        rhs and t_diagonal may be numpy arrays or cuda.DeviceArrays (HBM); for a
        device rhs, x is returned as a DeviceArray and nothing crosses PCIe but
        the solver's scalars.
        Raises SmhError when GMRES misses the tolerance unless with_stats is set,
        in which case (x, stats, converged) is returned."""
        b, b_ptr, x, x_ptr = _solve_buffers(rhs, self.shape[0])
        t, t_ptr = _device_or_host(t_diagonal)
        if tuple(t.shape) != (self.shape[0],):
            raise ValueError('t_diagonal and rhs must have the operator dimension %d' % self.shape[0])
        nrhs = 1 if b.ndim == 1 else b.shape[0]
        stats = np.zeros(8)
        mode = {'mixed': _lib.SOLVE_MIXED, 'mixed16': _lib.SOLVE_MIXED16, 'lu': _lib.SOLVE_LU}.get(
            precision, _lib.SOLVE_FP64)
        maxiter = 10 * self.shape[0] if maxiter is None else maxiter
        maxiter = int(min(maxiter, 2 ** 31 - 1))
        code = _lib.load().smh_solve(self._op, t_ptr, b_ptr, x_ptr, nrhs, float(tol), int(restart), maxiter, mode,
                                     _lib.dptr(stats))
        converged = code != _lib.SMH_ERR_NOCONV
        if not with_stats or converged:
            _lib.check(code, cu.context(), 'smh_solve')
        st = dict(iterations=int(stats[0]), matvecs_fp64=int(stats[1]), matvecs_fp32=int(stats[2]),
                  relative_residual=float(stats[3]), seconds=float(stats[4]), sweeps=int(stats[5]),
                  matvecs_fp16=int(stats[7]))
        return (x, st, converged) if with_stats else x

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


def _explicit_inputs(vacuum_wavelength, positions, layer_system, k_parallel, l_max, m_max):
    """Host arrays of the explicit coupling kernels for particles at `positions`:
    layer index among the occupied layers, each layer's wavenumber, heights
    above the layer's reference, direct coefficients and, unless the layer
    system is one medium, the per-layer-pair layer-mediated coefficients with
    the explicit quadrature's weights (its last contour interval left out)."""
    if isinstance(k_parallel, str) and k_parallel == 'default':
        k_parallel = flds.default_Sommerfeld_k_parallel_array
    layers = [layer_system.layer_number(p[2]) for p in positions]
    occupied = sorted(set(layers))
    lidx = {l: q for q, l in enumerate(occupied)}
    out = dict(
        layer=np.ascontiguousarray([lidx[l] for l in layers], dtype=np.int32),
        n_layers=len(occupied),
        k_layer=np.ascontiguousarray([complex(layer_system.wavenumber(l, vacuum_wavelength)) for l in occupied],
                                     dtype=np.complex128),
        z_rel=np.ascontiguousarray([p[2] - layer_system.reference_z(l) for p, l in zip(positions, layers)],
                                   dtype=np.float64),
        xyz=np.ascontiguousarray(positions, dtype=np.float64),
        direct_all=int(layer_system.is_degenerate()),
        coef=np.ascontiguousarray(look.direct_coefficients_volume(l_max, m_max)),
        kp=None, kz=None, lc=None, n_kappa=0)
    if not layer_system.is_degenerate():
        kp = np.ascontiguousarray(k_parallel, dtype=np.complex128)
        w = look.trapezoid_weights_drop_last(kp)
        nl = len(occupied)
        B = flds.blocksize(l_max, m_max)
        lc = np.empty((nl, nl, len(kp), 4, B, B), dtype=np.complex128)
        kz = np.empty((nl, len(kp)), dtype=np.complex128)
        for a, l1 in enumerate(occupied):
            for b, l2 in enumerate(occupied):
                lc[a, b], kz1, _ = look.layer_coefficients_pair(vacuum_wavelength, layer_system, kp, l1, l2, l_max,
                                                                 m_max, w)
            kz[a] = kz1
        out.update(kp=kp, kz=kz, lc=lc, n_kappa=len(kp))
    return out


def explicit_blocks(vacuum_wavelength, receivers, emitters, layer_system, k_parallel, l_max, m_max, x=None):
    """Explicit coupling blocks W[i, :, j, :] (receiver i, emitter j) of two
    position sets, padded to (l_max, m_max): complex128 [n_recv, B, n_emit, B],
    built on the GPU (direct_coupling_block + layer_mediated_coupling_block,
    direct_coupling.py:18-105, layer_mediated_coupling.py:215-340). With x
    ([n_emit, B]) the product W x ([n_recv, B]) is returned instead, formed on
    the device."""
    receivers = np.asarray(receivers, dtype=float).reshape(-1, 3)
    emitters = np.asarray(emitters, dtype=float).reshape(-1, 3)
    B = flds.blocksize(l_max, m_max)
    h = _explicit_inputs(vacuum_wavelength, np.concatenate([receivers, emitters]), layer_system, k_parallel, l_max,
                         m_max)
    if x is None:
        out = np.empty((len(receivers) * B, len(emitters) * B), dtype=np.complex128)
    else:
        x = np.ascontiguousarray(x, dtype=np.complex128).reshape(len(emitters) * B)
        out = np.empty(len(receivers) * B, dtype=np.complex128)
    ms = ctypes.c_double()
    ctx = cu.context()
    _lib.check(_lib.load().smh_explicit_blocks(ctx, len(receivers), len(emitters), _lib.dptr(h['xyz']), l_max, m_max,
                                               _lib.iptr(h['layer']), h['n_layers'], _lib.dptr(h['k_layer']),
                                               _lib.dptr(h['z_rel']), h['direct_all'], _lib.dptr(h['coef']),
                                               h['n_kappa'], _lib.dptr(h['kp']), _lib.dptr(h['kz']),
                                               _lib.dptr(h['lc']), _lib.dptr(x), _lib.dptr(out), ctypes.byref(ms)),
               ctx, 'smh_explicit_blocks')
    explicit_blocks.last_kernel_ms = ms.value
    if x is None:
        return out.reshape(len(receivers), B, len(emitters), B)
    return out.reshape(len(receivers), B)


explicit_blocks.last_kernel_ms = 0.0


class CouplingMatrixExplicit(_GpuCouplingOperator):
    """Explicit (dense) coupling matrix built on the MI355X; drop-in for
    smuthi's class of the same name (linear_system.py:441-486).

    Every block is direct_coupling_block (direct_coupling.py:18-105) plus
    layer_mediated_coupling_block (layer_mediated_coupling.py:215-340) for the
    pair's exact positions, in FP64, including pairs in different layers.
    The layer-mediated Sommerfeld integral uses the reference's quadrature,
    which leaves out the contour's last interval. The matrix stays on the
    device (complex128 plus a complex64 copy for FP32 matvecs);
    `coupling_matrix` downloads it in smuthi's system layout."""

    def __init__(self, vacuum_wavelength, particle_list, layer_system, k_parallel='default', use_pvwf_coupling=False,
                 pvwf_coupling_k_parallel=None):
        if use_pvwf_coupling:
            raise NotImplementedError('plane-wave mediated direct coupling is not built on the GPU path')
        self._init_common(particle_list)
        t0 = time.time()
        h = _explicit_inputs(vacuum_wavelength, [p.position for p in particle_list], layer_system, k_parallel,
                             self.l_max, self.m_max)
        self.host_seconds = time.time() - t0
        lib = _lib.load()
        ctx = cu.context()
        handle = ctypes.c_void_p()
        ms = ctypes.c_double()
        _lib.check(lib.smh_explicit_create(ctx, len(particle_list), _lib.dptr(h['xyz']), _lib.iptr(self._lmax),
                                           _lib.iptr(self._mmax), self.l_max, self.m_max, _lib.iptr(h['layer']),
                                           h['n_layers'], _lib.dptr(h['k_layer']), _lib.dptr(h['z_rel']),
                                           h['direct_all'], _lib.dptr(h['coef']), h['n_kappa'], _lib.dptr(h['kp']),
                                           _lib.dptr(h['kz']), _lib.dptr(h['lc']), ctypes.byref(ms),
                                           ctypes.byref(handle)), ctx, 'smh_explicit_create')
        self._op = handle
        self.build_kernel_ms = ms.value
        self.table_seconds = time.time() - t0
        self._make_linear_operator()

    @property
    def coupling_matrix(self):
        """The FP64 matrix [n, n] in smuthi's system layout (copied from the device on access)."""
        n_pad = len(self.particle_list) * self.blocksize
        full = np.empty((n_pad, n_pad), dtype=np.complex128)
        _lib.check(_lib.load().smh_op_download_table(self._op, _lib.dptr(full), None), cu.context(),
                   'smh_op_download_table')
        idx = np.concatenate([i * self.blocksize + np.array(
            [flds.multi_to_single_index(t, l, m, self.l_max, self.m_max)
             for t in range(2) for l in range(1, p.l_max + 1) for m in range(-min(l, p.m_max), min(l, p.m_max) + 1)])
            for i, p in enumerate(self.particle_list)])
        return full[np.ix_(idx, idx)]


class TMatrix(SystemMatrix):
    """Block-diagonal T of the particles (linear_system.py:410-438); spheres give a diagonal."""

    def __init__(self, particle_list):
        SystemMatrix.__init__(self, particle_list)
        diags = []
        for p in particle_list:
            t = np.asarray(p.t_matrix)
            d = np.diag(t)
            if np.count_nonzero(t - np.diag(d)):
                raise NotImplementedError('only diagonal (sphere) T-matrices are supported on the GPU path')
            diags.append(d)
        self.diagonal = np.ascontiguousarray(np.concatenate(diags), dtype=np.complex128)
        self.linear_operator = scipy.sparse.linalg.LinearOperator(
            shape=self.shape, matvec=lambda v: self.diagonal * v, matmat=lambda v: self.diagonal[:, None] * v,
            dtype=complex)

    def right_hand_side(self):
        a = np.concatenate([p.initial_field.coefficients for p in self.particle_list])
        return self.diagonal * a


class MasterMatrix(SystemMatrix):
    """M = 1 - T W as a LinearOperator (FP64 coupling matvec on the GPU)."""

    def __init__(self, t_matrix, coupling_matrix):
        SystemMatrix.__init__(self, t_matrix.particle_list)
        self.t_matrix = t_matrix
        self.coupling_matrix = coupling_matrix

        def apply(v):
            return v - self.t_matrix.diagonal * self.coupling_matrix.matvec(v, precision='double')

        self.linear_operator = scipy.sparse.linalg.LinearOperator(shape=self.shape, matvec=apply, dtype=complex)


class LinearSystem:
    """Assemble and solve (1 - T W) b = T a with smuthi's constructor arguments
    (linear_system.py:42-344). The coupling is a lookup operator or the
    explicit matrix, both built on the GPU; the solve runs on the GPU.
    Extra arguments: precision ('mixed': FP32 coupling matvecs inside FP64
    iterative refinement; 'mixed16': FP16-operand matvecs inside the same
    refinement, radial lookups; 'double': plain FP64 GMRES),
    cubic_interpolant, restart and maxiter. solver_type 'LU' factors the
    explicit matrix on the device (and, as the reference, refuses a lookup
    operator); periodicity and plane-wave mediated direct coupling raise
    NotImplementedError."""

    def __init__(self, particle_list, initial_field, layer_system, k_parallel='default', solver_type='LU',
                 solver_tolerance=1e-4, store_coupling_matrix=True, coupling_matrix_lookup_resolution=None,
                 interpolator_kind='cubic', cuda_blocksize=None, periodicity=None,
                 ewald_sum_separation_parameter='default', number_of_threads_periodic='default',
                 use_pvwf_coupling=False, pvwf_coupling_k_parallel=None, precision='mixed',
                 cubic_interpolant='spline', restart=20, maxiter=None):
        self.particle_list = particle_list
        self.initial_field = initial_field
        self.layer_system = layer_system
        self.k_parallel = k_parallel
        self.solver_type = solver_type
        self.solver_tolerance = solver_tolerance
        self.store_coupling_matrix = store_coupling_matrix
        self.coupling_matrix_lookup_resolution = coupling_matrix_lookup_resolution
        self.interpolator_kind = interpolator_kind
        self.cuda_blocksize = cu.default_blocksize if cuda_blocksize is None else cuda_blocksize
        self.precision = precision
        self.cubic_interpolant = cubic_interpolant
        self.restart = restart
        self.maxiter = maxiter
        if periodicity is not None:
            raise NotImplementedError('periodic particle arrangements are outside the GPU lookup path')
        self.periodicity = periodicity
        self.ewald_sum_separation_parameter = ewald_sum_separation_parameter
        self.number_of_threads_periodic = number_of_threads_periodic
        self.use_pvwf_coupling = use_pvwf_coupling
        self.pvwf_coupling_k_parallel = pvwf_coupling_k_parallel
        if use_pvwf_coupling:
            if coupling_matrix_lookup_resolution is None:
                raise NotImplementedError('plane-wave mediated direct coupling is not built on the GPU path')
            # the reference ignores the flag next to a lookup (linear_system.py:118-122)
            warnings.warn('The use_pvwf_coupling flag was defined in combination with a lookup. This combination '
                          'is not yet available. Ignoring the use_pvwf_coupling flag.')
        self.shape = SystemMatrix(particle_list).shape
        sys.stdout.write('Number of unknowns: %i\n' % self.shape[0])

    def prepare(self):
        self.compute_initial_field_coefficients()
        self.compute_t_matrix()
        self.compute_coupling_matrix()
        self.master_matrix = MasterMatrix(self.t_matrix, self.coupling_matrix)

    def compute_initial_field_coefficients(self):
        from .expansions import SphericalWaveExpansion
        a = self.initial_field.swe_coefficients(self.particle_list, self.layer_system)
        off = 0
        for p in self.particle_list:
            n = flds.blocksize(p.l_max, p.m_max)
            i = self.layer_system.layer_number(p.position[2])
            k = flds.angular_frequency(self.initial_field.vacuum_wavelength) * self.layer_system.refractive_indices[i]
            swe = SphericalWaveExpansion(k, p.l_max, p.m_max, 'regular', p.position,
                                         self.layer_system.lower_zlimit(i), self.layer_system.upper_zlimit(i))
            swe.coefficients = a[off:off + n]
            p.initial_field = swe
            off += n

    def compute_t_matrix(self):
        for p in self.particle_list:
            i = self.layer_system.layer_number(p.position[2])
            p.t_matrix = p.compute_t_matrix(self.initial_field.vacuum_wavelength,
                                            self.layer_system.refractive_indices[i])
        self.t_matrix = TMatrix(self.particle_list)

    def compute_coupling_matrix(self):
        """Lookup operator when a resolution is given and every particle sits in
        one layer, else the explicit matrix (linear_system.py:149-263)."""
        if not cu.use_gpu:
            raise RuntimeError('smuthi_amd provides the MI355X operators only: call '
                               'smuthi_amd.cuda.enable_gpu() first')
        if self.coupling_matrix_lookup_resolution is not None:
            layers = {self.layer_system.layer_number(p.position[2]) for p in self.particle_list}
            if len(layers) != 1:
                warnings.warn('Particles are not all in same layer. '
                              'Fall back to direct coupling matrix computation (no lookup).')
                self.coupling_matrix_lookup_resolution = None
            elif self.store_coupling_matrix:
                warnings.warn('Explicit matrix computation using lookup currently not implemented. '
                              'Disabling lookup.')
                self.coupling_matrix_lookup_resolution = None
        if self.coupling_matrix_lookup_resolution is None:
            if not self.store_coupling_matrix:
                warnings.warn('With lookup disabled, coupling matrix needs to be stored.')
                self.store_coupling_matrix = True
            self.coupling_matrix = CouplingMatrixExplicit(vacuum_wavelength=self.initial_field.vacuum_wavelength,
                                                          particle_list=self.particle_list,
                                                          layer_system=self.layer_system, k_parallel=self.k_parallel)
            return
        z = [p.position[2] for p in self.particle_list]
        kw = dict(vacuum_wavelength=self.initial_field.vacuum_wavelength, particle_list=self.particle_list,
                  layer_system=self.layer_system, k_parallel=self.k_parallel,
                  resolution=self.coupling_matrix_lookup_resolution, interpolator_kind=self.interpolator_kind)
        if z.count(z[0]) == len(z):
            self.coupling_matrix = CouplingMatrixRadialLookupCUDA(cuda_blocksize=self.cuda_blocksize,
                                                                  cubic_interpolant=self.cubic_interpolant, **kw)
        else:
            from .volume import CouplingMatrixVolumeLookupCUDA
            self.coupling_matrix = CouplingMatrixVolumeLookupCUDA(cuda_blocksize=self.cuda_blocksize,
                                                                  cubic_interpolant=self.cubic_interpolant, **kw)

    def solve(self, rhs=None):
        """Solve for the scattered-field coefficients; rhs may hold several right-hand sides [nrhs, n]."""
        from .expansions import SphericalWaveExpansion
        kind = self.solver_type.lower()
        if kind not in ('gmres', 'lgmres', 'lu'):
            raise ValueError('This solver type is currently not implemented.')
        single = rhs is None
        b = self.t_matrix.right_hand_side() if single else rhs
        tol, precision = self.solver_tolerance, self.precision
        if kind == 'lu':
            # linear_system.py:259-268: LU needs the stored (explicit) matrix; the
            # factors are kept on the device for later right-hand sides
            if not isinstance(self.coupling_matrix, CouplingMatrixExplicit):
                raise ValueError('LU factorization only possible with the option "store coupling matrix".')
            precision = 'lu'
        elif kind == 'lgmres':
            warnings.warn("solver_type 'lgmres' runs the GPU GMRES (same stopping rule)")
        x, self.solve_stats, converged = self.coupling_matrix.solve_master(
            self.t_matrix.diagonal, b, tol, precision=precision, restart=self.restart, maxiter=self.maxiter,
            with_stats=True)
        if not converged:
            warnings.warn('GMRES did not reach the tolerance: %s'
                          % _lib.load().smh_last_error(cu.context()).decode(errors='replace'))
        if not single:
            return x
        off = 0
        for p in self.particle_list:
            n = flds.blocksize(p.l_max, p.m_max)
            i = self.layer_system.layer_number(p.position[2])
            k = flds.angular_frequency(self.initial_field.vacuum_wavelength) * self.layer_system.refractive_indices[i]
            swe = SphericalWaveExpansion(k, p.l_max, p.m_max, 'outgoing', p.position,
                                         self.layer_system.lower_zlimit(i), self.layer_system.upper_zlimit(i))
            swe.coefficients = x[off:off + n].copy()
            swe.validity_conditions.append(p.is_outside)
            p.scattered_field = swe
            off += n
        return x
