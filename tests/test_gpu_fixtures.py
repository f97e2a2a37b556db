"""The GPU path against the reference's system-test literals, through the
drop-in Simulation with the reference's arguments (solver_type 'LU' and the
stored explicit coupling matrix, the reference's defaults).

* Dipole-illuminated grid of 21 spheres, l_max 4..7 in one medium
  (tests/system_tests/test_direct_radial_and_volume_particle_coupling.py:26-84):
  all 1748 coefficients of the '2D' and '3D' layouts against
  tests/golden/dipole_grid.npz under the reference's element check.
* Two spheres in a slab waveguide (tests/system_tests/test_two_spheres_in_waveguide.py:14-82):
  the split-layer system gives the same coupling matrix (1e-3), four
  coefficients to 1e-4, top / bottom scattered power to 1e-3.
"""
import os

import numpy as np
import pytest

from test_oracle_fixtures import WG_BOTTOM, WG_LITERALS, WG_TOP, WG_WL, dipole_grid, is_same

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


@pytest.mark.parametrize('dimension', ['2D', '3D'])
def test_dipole_grid_matches_golden(gpu, dimension):
    from smuthi_amd.initial_field import DipoleSource
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.simulation import Simulation
    print('dipole grid', dimension, flush=True)
    spec, dpos = dipole_grid(dimension)
    parts = [Sphere(position=p, refractive_index=5 + 0.5j, radius=120, l_max=l) for p, l in spec]
    sim = Simulation(layer_system=LayerSystem(thicknesses=[0, 0], refractive_indices=[1, 1]), particle_list=parts,
                     initial_field=DipoleSource(vacuum_wavelength=416, dipole_moment=[1, 1, 1], position=dpos),
                     solver_type='LU', store_coupling_matrix=True, log_to_terminal=False)
    sim.run()
    got = np.concatenate([p.scattered_field.coefficients for p in sim.particle_list])
    want = np.load(os.path.join(GOLDEN, 'dipole_grid.npz'))['known_%s' % dimension.lower()]
    assert got.shape == want.shape
    bad = [i for i in range(len(want)) if not is_same(want[i], got[i])]
    worst = np.max(np.abs(got - want) / np.abs(want))
    print('  largest relative deviation %.2e, %d of %d outside 1e-7' % (worst, len(bad), len(want)), flush=True)
    assert not bad, bad[:5]


def _waveguide(thick, nidx):
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.simulation import Simulation
    parts = [Sphere(position=[100, 100, 150], refractive_index=2.4 + 0.0j, radius=120, l_max=3),
             Sphere(position=[-100, -100, 250], refractive_index=1.9 + 0.1j, radius=120, l_max=3)]
    pw = PlaneWave(vacuum_wavelength=WG_WL, polar_angle=np.pi, azimuthal_angle=0, polarization=0, amplitude=1,
                   reference_point=[0, 0, 400])
    sim = Simulation(layer_system=LayerSystem(thick, nidx), particle_list=parts, initial_field=pw,
                     neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 20], neff_resolution=1e-3,
                     log_to_terminal=False)
    sim.run()
    return sim


def test_two_spheres_in_waveguide(gpu):
    from smuthi_amd import far_field as farf
    print('two spheres in a waveguide', flush=True)
    sim1 = _waveguide([0, 400, 0], [1.5, 1.7, 1])
    sim2 = _waveguide([0, 200, 200, 0], [1.5, 1.7, 1.7, 1])
    A1 = sim1.linear_system.coupling_matrix.coupling_matrix
    A2 = sim2.linear_system.coupling_matrix.coupling_matrix
    rel = np.linalg.norm(A1 - A2) / np.linalg.norm(A1)
    print('  equivalent layer systems: %.2e' % rel, flush=True)
    assert rel < 1e-3
    b = np.concatenate([p.scattered_field.coefficients for p in sim1.particle_list])
    for k, v in WG_LITERALS.items():
        print('  b[%d] = %s, relative deviation %.2e' % (k, b[k], abs((b[k] - v) / v)), flush=True)
        assert abs((b[k] - v) / v) < 1e-4, k
    res = np.pi / 360
    ff = farf.scattered_far_field(WG_WL, sim1.particle_list, sim1.layer_system,
                                  polar_angles=np.arange(0, np.pi + 0.5 * res, res),
                                  azimuthal_angles=np.arange(0, 2 * np.pi + 0.5 * res, res))
    top, bottom = float(np.sum(ff.top().integral())), float(np.sum(ff.bottom().integral()))
    print('  power top %.6e, bottom %.6e' % (top, bottom), flush=True)
    assert abs(top - WG_TOP) < 1e-3 * WG_TOP
    assert abs(bottom - WG_BOTTOM) < 1e-3 * WG_BOTTOM
