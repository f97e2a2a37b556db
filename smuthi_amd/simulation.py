"""Simulation driver with smuthi's interface (simulation.py:92-414)."""
import sys
import time

import numpy as np

from . import fields as flds
from . import lookup as look
from .linear_system import LinearSystem


class Simulation:
    """Set up contours, prepare and solve the linear system; run() returns
    (preparation, solution, post-processing) seconds like smuthi."""

    def __init__(self, layer_system=None, particle_list=None, initial_field=None, k_parallel='default',
                 angular_resolution=np.pi / 360, neff_waypoints=None, neff_imag=1e-2, neff_max=None,
                 neff_max_offset=1, neff_resolution=1e-2, neff_minimal_branchpoint_distance=None,
                 overwrite_default_contours=True, solver_type='LU', solver_tolerance=1e-4,
                 store_coupling_matrix=True, coupling_matrix_lookup_resolution=None,
                 coupling_matrix_interpolator_kind='linear', length_unit='length unit', input_file=None,
                 output_dir='smuthi_output', save_after_run=False, log_to_file=False, log_to_terminal=True,
                 check_circumscribing_spheres=True, do_sanity_check=True, periodicity=None,
                 ewald_sum_separation_parameter='default', number_of_threads_periodic='default',
                 use_pvwf_coupling=False, pvwf_coupling_neff_max=None, pvwf_coupling_neff_resolution=1e-2,
                 precision='mixed', cubic_interpolant='spline', restart=20, maxiter=None):
        # arguments of the reference's constructor (simulation.py:92-123) that
        # select paths this package does not build raise instead of being dropped
        if periodicity is not None:
            raise NotImplementedError('periodic particle arrangements are outside the GPU lookup path')
        if use_pvwf_coupling and coupling_matrix_lookup_resolution is None:
            raise NotImplementedError('plane-wave mediated direct coupling is not built on the GPU path')
        if save_after_run or log_to_file:
            raise NotImplementedError('saving the simulation and logging to a file are not part of this package')
        self.length_unit = length_unit
        self.input_file = input_file
        self.output_dir = output_dir
        self.save_after_run = save_after_run
        self.log_to_file = log_to_file
        self.log_to_terminal = log_to_terminal
        self.check_circumscribing_spheres = check_circumscribing_spheres
        self.periodicity = periodicity
        self.ewald_sum_separation_parameter = ewald_sum_separation_parameter
        self.number_of_threads_periodic = number_of_threads_periodic
        self.use_pvwf_coupling = use_pvwf_coupling
        self.pvwf_coupling_neff_max = pvwf_coupling_neff_max
        self.pvwf_coupling_neff_resolution = pvwf_coupling_neff_resolution
        self.layer_system = layer_system
        self.particle_list = particle_list
        self.initial_field = initial_field
        self.k_parallel = k_parallel
        self.angular_resolution = angular_resolution
        self.neff_waypoints = neff_waypoints
        self.neff_imag = neff_imag
        self.neff_max = neff_max
        self.neff_max_offset = neff_max_offset
        self.neff_resolution = neff_resolution
        self.neff_minimal_branchpoint_distance = neff_minimal_branchpoint_distance
        self.overwrite_default_contours = overwrite_default_contours
        self.solver_type = solver_type
        self.solver_tolerance = solver_tolerance
        self.store_coupling_matrix = store_coupling_matrix
        self.coupling_matrix_lookup_resolution = coupling_matrix_lookup_resolution
        self.coupling_matrix_interpolator_kind = coupling_matrix_interpolator_kind
        self.do_sanity_check = do_sanity_check
        self.precision = precision
        self.cubic_interpolant = cubic_interpolant
        self.restart = restart
        self.maxiter = maxiter
        self.post_processing = None

    def largest_lateral_distance(self):
        xy = np.array([p.position[:2] for p in self.particle_list], dtype=float)
        return look.pair_distance_extremes(xy)[0] if len(xy) > 1 else 0.0

    def sanity_check(self):
        maxrho = self.largest_lateral_distance()
        k = flds.angular_frequency(self.initial_field.vacuum_wavelength)
        thresh = 2 / (maxrho * k)
        if self.neff_resolution is not None and isinstance(self.k_parallel, str) and self.neff_resolution > thresh:
            sys.stdout.write('Warning: Sanity check for neff_resolution failed! Recommended < %.3e\n' % thresh)

    def set_default_contours(self):
        flds.default_Sommerfeld_k_parallel_array = flds.reasonable_Sommerfeld_kpar_contour(
            vacuum_wavelength=self.initial_field.vacuum_wavelength, neff_waypoints=self.neff_waypoints,
            layer_refractive_indices=self.layer_system.refractive_indices, neff_imag=self.neff_imag,
            neff_max=self.neff_max, neff_max_offset=self.neff_max_offset, neff_resolution=self.neff_resolution,
            neff_minimal_branchpoint_distance=self.neff_minimal_branchpoint_distance)
        flds.default_initial_field_k_parallel_array = flds.default_Sommerfeld_k_parallel_array
        if flds.default_azimuthal_angles is None:
            flds.default_azimuthal_angles, flds.default_polar_angles = flds.angular_arrays(self.angular_resolution)

    def initialize_linear_system(self):
        self.linear_system = LinearSystem(
            particle_list=self.particle_list, initial_field=self.initial_field, layer_system=self.layer_system,
            k_parallel=self.k_parallel, solver_type=self.solver_type, solver_tolerance=self.solver_tolerance,
            store_coupling_matrix=self.store_coupling_matrix,
            coupling_matrix_lookup_resolution=self.coupling_matrix_lookup_resolution,
            interpolator_kind=self.coupling_matrix_interpolator_kind, periodicity=self.periodicity,
            ewald_sum_separation_parameter=self.ewald_sum_separation_parameter,
            number_of_threads_periodic=self.number_of_threads_periodic, use_pvwf_coupling=self.use_pvwf_coupling,
            precision=self.precision,
            cubic_interpolant=self.cubic_interpolant, restart=self.restart, maxiter=self.maxiter)

    def run(self):
        if self.do_sanity_check and len(self.particle_list) > 1:
            self.sanity_check()
        if flds.default_Sommerfeld_k_parallel_array is None or self.overwrite_default_contours:
            self.set_default_contours()
        t0 = time.time()
        self.initialize_linear_system()
        self.linear_system.prepare()
        t1 = time.time()
        self.linear_system.solve()
        t2 = time.time()
        if self.post_processing:
            self.post_processing.run(self)
        t3 = time.time()
        return t1 - t0, t2 - t1, t3 - t2
