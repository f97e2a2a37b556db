"""Drop-in argument handling of Simulation / LinearSystem (no device needed):
the reference's defaults and the reference arguments that select paths this
package does not build."""
import inspect

import numpy as np
import pytest

from smuthi_amd.layers import LayerSystem
from smuthi_amd.linear_system import LinearSystem
from smuthi_amd.particles import Sphere
from smuthi_amd.simulation import Simulation


def _parts():
    return [Sphere([0, 0, 200], 1.5 + 0.01j, 100, 3, 3), Sphere([300, 0, 200], 1.5 + 0.01j, 100, 3, 3)]


def test_linear_system_defaults_match_the_reference():
    # linear_system.py:80-82: solver_type='LU', store_coupling_matrix=True, interpolator_kind='cubic'
    sig = inspect.signature(LinearSystem.__init__).parameters
    assert sig['solver_type'].default == 'LU'
    assert sig['store_coupling_matrix'].default is True
    assert sig['interpolator_kind'].default == 'cubic'
    assert sig['solver_tolerance'].default == 1e-4
    # simulation.py:104-109
    sig = inspect.signature(Simulation.__init__).parameters
    assert sig['solver_type'].default == 'LU'
    assert sig['store_coupling_matrix'].default is True
    assert sig['coupling_matrix_interpolator_kind'].default == 'linear'


def test_unbuilt_paths_raise_instead_of_being_dropped():
    ls = LayerSystem([0, 0], [1.52, 1])
    with pytest.raises(NotImplementedError):
        Simulation(layer_system=ls, particle_list=_parts(), periodicity=(np.array([1.0, 0, 0]), np.array([0, 1.0, 0])))
    with pytest.raises(NotImplementedError):
        Simulation(layer_system=ls, particle_list=_parts(), use_pvwf_coupling=True)
    with pytest.raises(NotImplementedError):
        Simulation(layer_system=ls, particle_list=_parts(), save_after_run=True)
    with pytest.raises(TypeError):
        Simulation(layer_system=ls, particle_list=_parts(), no_such_argument=1)
    with pytest.raises(NotImplementedError):
        LinearSystem(_parts(), None, ls, periodicity=(np.array([1.0, 0, 0]), np.array([0, 1.0, 0])))
    with pytest.raises(NotImplementedError):
        LinearSystem(_parts(), None, ls, use_pvwf_coupling=True)
    with pytest.raises(TypeError):
        LinearSystem(_parts(), None, ls, no_such_argument=1)
    # next to a lookup the reference warns and ignores the flag (linear_system.py:118-122)
    with pytest.warns(UserWarning):
        LinearSystem(_parts(), None, ls, use_pvwf_coupling=True, coupling_matrix_lookup_resolution=5.0)


def test_reference_arguments_are_accepted():
    ls = LayerSystem([0, 0], [1.52, 1])
    sim = Simulation(layer_system=ls, particle_list=_parts(), length_unit='nm', output_dir='out', log_to_terminal=False,
                     check_circumscribing_spheres=False, ewald_sum_separation_parameter='default',
                     number_of_threads_periodic='default', pvwf_coupling_neff_max=None,
                     pvwf_coupling_neff_resolution=1e-2)
    assert sim.length_unit == 'nm' and sim.periodicity is None
