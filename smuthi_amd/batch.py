"""Many particle ensembles and many incident fields solved together on the GPU.

smuthi solves one ensemble for one initial field per Simulation
(simulation.py:363-414) and rebuilds its lookup table each time
(linear_system.py:183-225). EnsembleBatch puts N ensembles that live in one
layer system into one device operator: they share one lookup table (sized
for the largest), every coupling matvec is one set of launches over the pairs
of all ensembles (pairs couple particles of one ensemble only), and each
(ensemble, initial field) pair is its own GMRES system with its own Arnoldi
scalars and convergence, kBatch fields at a time
(smh_radial_create_ensembles / smh_volume_create_ensembles, smh_solve). Every
returned solution meets the FP64 stopping rule of the reference's GMRES,
||T a - M b|| <= tol ||T a|| (linear_system.py:271-286), per ensemble and
field.
"""
import ctypes
import time

import numpy as np

from . import _lib
from . import cuda as cu
from . import fields as flds
from . import lookup as look
from .linear_system import TMatrix, _GpuCouplingOperator, _interp_code as _radial_interp
from .volume import _interp_code as _volume_interp


class _EnsembleOperator(_GpuCouplingOperator):
    """The device operator over all ensembles (system vectors concatenate the
    ensembles' vectors)."""


class EnsembleBatch:
    """ensembles: list of particle lists (spheres), all particles of all
    ensembles in one layer of layer_system. The lookup is radial when every
    particle of an ensemble shares one height and all ensembles use the same
    height, else volume. interpolator_kind / cubic_interpolant / resolution as
    for CouplingMatrixVolumeLookupCUDA (default resolution: wavelength / 100).
    Attributes: operator (the device operator, with matvec / solve_master),
    offsets (first unknown of every ensemble and the total), t_diagonal
    (T-matrix diagonal of all particles), table_seconds."""

    def __init__(self, vacuum_wavelength, ensembles, layer_system, k_parallel='default', resolution=None,
                 interpolator_kind='cubic', cubic_interpolant='spline'):
        if not ensembles or any(len(e) == 0 for e in ensembles):
            raise ValueError('every ensemble needs at least one particle')
        self.vacuum_wavelength = vacuum_wavelength
        self.ensembles = [list(e) for e in ensembles]
        self.layer_system = layer_system
        parts = [p for e in self.ensembles for p in e]
        self.particle_list = parts
        layers = {layer_system.layer_number(p.position[2]) for p in parts}
        if len(layers) != 1:
            raise ValueError('all particles of all ensembles must lie in one layer')
        i_s = layers.pop()
        self.resolution = vacuum_wavelength / 100 if resolution is None else resolution
        if isinstance(k_parallel, str) and k_parallel == 'default':
            k_parallel = flds.default_Sommerfeld_k_parallel_array
        self.k_parallel = k_parallel
        n_part = np.array([len(e) for e in self.ensembles], dtype=np.int64)
        self.particle_offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(n_part)]), dtype=np.int64)
        sizes = np.array([flds.blocksize(p.l_max, p.m_max) for p in parts], dtype=np.int64)
        unk = np.concatenate([[0], np.cumsum(sizes)])
        self.offsets = unk[self.particle_offsets]
        xyz = np.ascontiguousarray([p.position for p in parts], dtype=np.float64)
        # the table covers every ensemble: largest in-plane distance and smallest
        # pair distance within an ensemble, z range of all particles
        rho_max, rho_min_pair = 0.0, np.inf
        for e0, e1 in zip(self.particle_offsets[:-1], self.particle_offsets[1:]):
            if e1 - e0 >= 2:
                a, b = look.pair_distance_extremes(xyz[e0:e1, :2])
                rho_max, rho_min_pair = max(rho_max, a), min(rho_min_pair, b)
        if not np.isfinite(rho_min_pair):
            rho_max, rho_min_pair = self.resolution, self.resolution
        op = _EnsembleOperator.__new__(_EnsembleOperator)
        op._init_common(parts)
        self.operator = op
        radial = len(set(xyz[:, 2])) == 1
        self.kind = 'radial' if radial else 'volume'
        lib = _lib.load()
        ctx = cu.context()
        handle = ctypes.c_void_p()
        k_is = complex(layer_system.wavenumber(i_s, vacuum_wavelength))
        t0 = time.time()
        if radial:
            interp = _radial_interp(interpolator_kind, cubic_interpolant)
            rho = look.radial_grid(rho_max, self.resolution)
            self.radial_distance_array = rho
            xy = np.ascontiguousarray(xyz[:, :2])
            _lib.check(lib.smh_radial_create_ensembles(ctx, len(self.ensembles), _lib.i64ptr(self.particle_offsets),
                                                       _lib.dptr(xy), _lib.iptr(op._lmax), _lib.iptr(op._mmax),
                                                       op.l_max, op.m_max, float(rho[0]), float(self.resolution),
                                                       len(rho), interp, ctypes.byref(handle)), ctx,
                       'smh_radial_create_ensembles')
            op._op = handle
            dz = xyz[0, 2] - layer_system.reference_z(i_s)
            coef = np.ascontiguousarray(look.direct_coefficients_radial(op.l_max, op.m_max))
            kp = lm = None
            if not layer_system.is_degenerate():
                kp = np.ascontiguousarray(k_parallel, dtype=np.complex128)
                lm = np.ascontiguousarray(look.layer_coefficients_radial(vacuum_wavelength, layer_system, kp, i_s, dz,
                                                                         op.l_max, op.m_max))
            _lib.check(lib.smh_radial_build_table(op._op, k_is.real, k_is.imag, rho_min_pair / 2, _lib.dptr(coef),
                                                  0 if kp is None else len(kp), _lib.dptr(kp), _lib.dptr(lm)), ctx,
                       'smh_radial_build_table')
        else:
            interp = _volume_interp(interpolator_kind, cubic_interpolant)
            rho, sz, dz = look.volume_grids_from_extremes(rho_max, xyz[:, 2].min(), xyz[:, 2].max(), self.resolution)
            self.rho_array, self.sum_z_array, self.diff_z_array = rho, sz, dz
            _lib.check(lib.smh_volume_create_ensembles(ctx, len(self.ensembles), _lib.i64ptr(self.particle_offsets),
                                                       _lib.dptr(xyz), _lib.iptr(op._lmax), _lib.iptr(op._mmax),
                                                       op.l_max, op.m_max, float(rho[0]), float(sz[0]), float(dz[0]),
                                                       float(self.resolution), len(rho), len(sz), interp,
                                                       ctypes.byref(handle)), ctx, 'smh_volume_create_ensembles')
            op._op = handle
            coef = np.ascontiguousarray(look.direct_coefficients_volume(op.l_max, op.m_max))
            kp = kz = lc = None
            n_k = 0
            if not layer_system.is_degenerate():
                kp = np.ascontiguousarray(k_parallel, dtype=np.complex128)
                lc, kz = look.layer_coefficients_volume(vacuum_wavelength, layer_system, kp, i_s, op.l_max, op.m_max)
                lc, kz = np.ascontiguousarray(lc), np.ascontiguousarray(kz, dtype=np.complex128)
                n_k = len(kp)
            _lib.check(lib.smh_volume_build_table(op._op, k_is.real, k_is.imag, rho_min_pair / 2,
                                                  float(layer_system.reference_z(i_s)), _lib.dptr(coef), n_k,
                                                  _lib.dptr(kp), _lib.dptr(kz), _lib.dptr(lc)), ctx,
                       'smh_volume_build_table')
        self.table_seconds = time.time() - t0
        n_medium = layer_system.refractive_indices[i_s]
        for p in parts:
            p.t_matrix = p.compute_t_matrix(vacuum_wavelength, n_medium)
        self.t_matrix = TMatrix(parts)
        self.t_diagonal = self.t_matrix.diagonal
        op._make_linear_operator()

    @property
    def n_unknowns(self):
        return int(self.offsets[-1])

    def right_hand_sides(self, initial_fields):
        """T a of every ensemble for every initial field: [n_fields, n_unknowns]."""
        out = np.empty((len(initial_fields), self.n_unknowns), dtype=np.complex128)
        for j, f in enumerate(initial_fields):
            out[j] = self.t_diagonal * f.swe_coefficients(self.particle_list, self.layer_system)
        return out

    def split(self, x):
        """Per-ensemble parts of [..., n_unknowns] vectors."""
        return [x[..., a:b] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def set_preconditioner(self, blocks, neighbours=None):
        """Right preconditioner for the solves: block-diagonal blocks [n_particles, B, B], or
        block-sparse blocks [n_particles, k1, B, B] with neighbours [n_particles, k1] (emitter of each
        block, -1 empty), on the padded multipole index of blocksize(l_max, m_max); None removes it."""
        return set_preconditioner(self.operator, blocks, neighbours)

    def solve(self, rhs, tol=1e-6, precision='mixed', restart=20, maxiter=None):
        """Solve every ensemble for every right-hand side ([n_fields, n_unknowns], e.g. from
        right_hand_sides): returns x [n_fields, n_unknowns]. Raises SmhError when a system misses
        the tolerance."""
        return self.operator.solve_master(self.t_diagonal, rhs, tol, precision=precision, restart=restart,
                                          maxiter=maxiter)

    def solve_with_stats(self, rhs, tol=1e-6, precision='mixed', restart=20, maxiter=None):
        """As solve, returning (x, stats, converged) instead of raising."""
        return self.operator.solve_master(self.t_diagonal, rhs, tol, precision=precision, restart=restart,
                                          maxiter=maxiter, with_stats=True)


def set_preconditioner(operator, blocks, neighbours=None):
    """smh_op_set_preconditioner on a device operator (see EnsembleBatch.set_preconditioner)."""
    if blocks is None:
        _lib.check(_lib.load().smh_op_set_preconditioner(operator._op, 0, None, None), cu.context(),
                   'smh_op_set_preconditioner')
        return
    b = np.asarray(blocks, dtype=np.complex128)
    n, B = len(operator.particle_list), operator.blocksize
    if neighbours is None:
        if b.shape != (n, B, B):
            raise ValueError('block-diagonal preconditioner blocks must be [n_particles, B, B]')
        b = b[:, None]
        neighbours = np.arange(n, dtype=np.int32)[:, None]
    nb = np.ascontiguousarray(neighbours, dtype=np.int32)
    if b.ndim != 4 or b.shape[0] != n or b.shape[2:] != (B, B) or nb.shape != b.shape[:2]:
        raise ValueError('preconditioner blocks must be [n_particles, k1, B, B] with neighbours [n_particles, k1]')
    b = np.ascontiguousarray(b)
    _lib.check(_lib.load().smh_op_set_preconditioner(operator._op, nb.shape[1], _lib.iptr(nb), _lib.dptr(b)),
               cu.context(), 'smh_op_set_preconditioner')
