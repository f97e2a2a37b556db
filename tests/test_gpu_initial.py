"""GPU parity of the non-plane-wave right-hand sides against the FP64 oracle
and the reference's golden vectors: Gaussian beam SWE coefficients
(GaussianBeam, initial_field.py:157-233, converted by the PWE-to-SWE kernel,
transformations.py:77-130) and dipole sources (DipoleSource /
DipoleCollection, initial_field.py:331-709, through the explicit coupling
blocks)."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import initial as oin
from oracle import layers as ol
from oracle import system as osys

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


class _P:
    def __init__(self, position, l_max=3, m_max=3):
        self.position = list(position)
        self.l_max = l_max
        self.m_max = m_max


def test_gaussian_beam_golden_vector(gpu):
    # tests/unit_tests/initial_field_tests/test_initial_field_gauss.py:7-35
    from smuthi_amd import fields as flds
    from smuthi_amd.initial_field import GaussianBeam
    from smuthi_amd.layers import LayerSystem
    ld = 532
    kp = np.linspace(0, 0.99999, 1000) * flds.angular_frequency(ld)
    al = np.linspace(0, 2 * np.pi, 1000)
    print('gaussian beam golden vector', flush=True)
    beam = GaussianBeam(ld, 0, 0.2 * np.pi, 1, 4000, k_parallel_array=kp, azimuthal_angles_array=al, amplitude=1,
                        reference_point=[-100, 100, 200])
    a = beam.swe_coefficients([_P([100, -200, 300])], LayerSystem([0, 0], [1, 1]))
    for idx, want in ((0, 0.4040275 - 1.6689055j), (9, -0.0115840 + 0.0107125j),
                      (20, -3.1855342e-04 - 7.7089434e-04j)):
        assert abs((a[idx] - want) / want) < 1e-5, idx
    beam_o = oin.GaussianBeam(ld, 0, 0.2 * np.pi, 1, 4000, kp, al, amplitude=1, reference_point=[-100, 100, 200])
    want = osys.initial_field_coefficients(beam_o, osys.Sphere([100, -200, 300], 1.5, 100, 3, 3),
                                           ol.LayerSystem([0, 0], [1, 1]))
    assert np.abs(a - want).max() < 1e-11 * np.abs(want).max()


def test_gaussian_beam_many_particles_layered(gpu):
    # beam from the top through a film, complex Sommerfeld contour, particles in all three layers with
    # different cut-offs, against the oracle's pwe_to_swe_conversion per particle
    from smuthi_amd.initial_field import GaussianBeam
    from smuthi_amd.layers import LayerSystem
    wl = 550.0
    thick, nidx = [0, 250, 0], [1.5, 1.9 + 0.02j, 1.0]
    kp = ob.reasonable_Sommerfeld_kpar_contour(wl, layer_refractive_indices=[1.5, 1.9, 1.0], neff_resolution=5e-3)
    kp = kp[np.real(kp) < 0.95 * 2 * np.pi / wl]  # propagating part of the top medium
    al = np.linspace(0, 2 * np.pi, 181)
    rng = np.random.default_rng(3)
    parts = [_P([rng.uniform(-400, 400), rng.uniform(-400, 400), z], l, m)
             for z, l, m in ((-120, 3, 3), (-60, 2, 2), (80, 3, 2), (170, 3, 3), (330, 2, 1), (500, 3, 3))]
    print('gaussian beam, 6 particles in 3 layers', flush=True)
    beam = GaussianBeam(wl, np.pi - 0.35, 1.1, 1, 1200.0, k_parallel_array=kp, azimuthal_angles_array=al,
                        amplitude=0.8, reference_point=[50, -20, 300])
    got = beam.swe_coefficients(parts, LayerSystem(thick, nidx))
    beam_o = oin.GaussianBeam(wl, np.pi - 0.35, 1.1, 1, 1200.0, kp, al, amplitude=0.8, reference_point=[50, -20, 300])
    want = np.concatenate([osys.initial_field_coefficients(beam_o, osys.Sphere(p.position, 1.5, 50, p.l_max, p.m_max),
                                                           ol.LayerSystem(thick, nidx)) for p in parts])
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_gaussian_beam_high_orders(gpu):
    # l_max 5..7 and m_max 0: the runtime-order PWE-to-SWE kernel (orders m in chunks of 5 per launch)
    from smuthi_amd.initial_field import GaussianBeam
    from smuthi_amd.layers import LayerSystem
    wl = 550.0
    thick, nidx = [0, 250, 0], [1.5, 1.9 + 0.02j, 1.0]
    kp = ob.reasonable_Sommerfeld_kpar_contour(wl, layer_refractive_indices=[1.5, 1.9, 1.0], neff_resolution=5e-3)
    kp = kp[np.real(kp) < 0.95 * 2 * np.pi / wl]
    al = np.linspace(0, 2 * np.pi, 181)
    rng = np.random.default_rng(5)
    parts = [_P([rng.uniform(-400, 400), rng.uniform(-400, 400), z], l, m)
             for z, l, m in ((-90, 6, 6), (100, 5, 3), (200, 7, 2), (420, 5, 5), (460, 3, 0))]
    beam = GaussianBeam(wl, np.pi - 0.35, 1.1, 1, 1200.0, k_parallel_array=kp, azimuthal_angles_array=al,
                        amplitude=0.8, reference_point=[50, -20, 300])
    got = beam.swe_coefficients(parts, LayerSystem(thick, nidx))
    beam_o = oin.GaussianBeam(wl, np.pi - 0.35, 1.1, 1, 1200.0, kp, al, amplitude=0.8, reference_point=[50, -20, 300])
    want = np.concatenate([osys.initial_field_coefficients(beam_o, osys.Sphere(p.position, 1.5, 50, p.l_max, p.m_max),
                                                           ol.LayerSystem(thick, nidx)) for p in parts])
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_dipole_golden_vector(gpu):
    # tests/unit_tests/initial_field_tests/test_initial_field_dipole.py:8-52
    from smuthi_amd.initial_field import DipoleSource
    from smuthi_amd.layers import LayerSystem
    ld = 550
    kp = ob.reasonable_Sommerfeld_kpar_contour(vacuum_wavelength=ld,
                                               neff_waypoints=[0, 0.8, 0.8 - 0.1j, 2.1 - 0.1j, 2.1, 4],
                                               neff_resolution=1e-2)
    thick, n = [0, 200, 200, 0], [1, 1.5, 2 + 1e-2j, 1 + 5j]
    print('dipole golden vector', flush=True)
    dip = DipoleSource(ld, [1e7, 2e7, 3e7], [300, -200, 100], k_parallel_array=kp)
    a = dip.swe_coefficients([_P([100, 200, 300])], LayerSystem(thick, n))
    for idx, want in ((0, -0.042563330382109 - 1.073039889335632j), (10, 0.496704638004303 + 0.908744692802429j),
                      (20, 1.209331750869751 - 2.165398836135864j), (29, -0.801326990127563 - 0.706078171730042j)):
        assert abs((a[idx] - want) / want) < 1e-4, idx
    want = oin.DipoleSource(ld, [1e7, 2e7, 3e7], [300, -200, 100], kp).spherical_wave_expansion(
        osys.Sphere([100, 200, 300], 1.5, 100, 3, 3), ol.LayerSystem(thick, n))
    assert np.abs(a - want).max() < 1e-10 * np.abs(want).max()


def test_dipole_collection_matches_oracle(gpu):
    # tests/unit_tests/initial_field_tests/test_initial_field_dipole_array.py:8-46 geometry: two dipoles, three
    # spheres, one sphere in a dipole's layer (direct part) and two across the film
    from smuthi_amd.initial_field import DipoleCollection, DipoleSource
    from smuthi_amd.layers import LayerSystem
    ld = 550
    kp = ob.reasonable_Sommerfeld_kpar_contour(vacuum_wavelength=ld, neff_waypoints=[0, 0.8, 0.8 - 0.1j, 2.1 - 0.1j, 2.1, 3],
                                               neff_resolution=5e-3)
    thick, n = [0, 400, 0], [1, 2, 1.5]
    dips = [([100, -100, 100], [1e7, 2e7, 3e7]), ([-100, 100, -100], [-2e7, 3e7, 1e7])]
    spheres = [([200, 200, 500], 3, 3), ([200, -200, 500], 3, 2), ([-200, -200, -400], 2, 2)]
    coll = DipoleCollection(ld, k_parallel_array=kp)
    for pos, mom in dips:
        coll.append(DipoleSource(ld, mom, pos, k_parallel_array=kp))
    print('dipole collection', flush=True)
    got = coll.swe_coefficients([_P(p, l, m) for p, l, m in spheres], LayerSystem(thick, n))
    oc = oin.DipoleCollection([oin.DipoleSource(ld, mom, pos, kp) for pos, mom in dips])
    want = np.concatenate([oc.spherical_wave_expansion(osys.Sphere(p, 2.4, 100, l, m), ol.LayerSystem(thick, n))
                           for p, l, m in spheres])
    assert np.abs(got - want).max() < 1e-10 * np.abs(want).max()


def test_solve_with_gaussian_beam(gpu):
    # the lookup solve with a Gaussian beam right-hand side against the oracle's LU of the same operator
    from oracle import coupling as oc
    from smuthi_amd.initial_field import GaussianBeam
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.simulation import Simulation
    wl = 550.0
    kp = ob.reasonable_Sommerfeld_kpar_contour(wl, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    al = np.linspace(0, 2 * np.pi, 121)
    plist = [osys.Sphere([0, 0, 150], 2.0 + 0.05j, 100, 3, 3), osys.Sphere([260, -40, 150], 1.7, 90, 3, 3),
             osys.Sphere([60, 250, 150], 1.9 + 0.01j, 80, 3, 3)]
    ls_args = ([0, 0], [1.52, 1])
    beam_args = (wl, np.pi, 0.0, 0, 900.0)
    kpb = kp[np.real(kp) < 0.99 * 2 * np.pi / wl]
    pparts = [Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in plist]
    sim = Simulation(layer_system=LayerSystem(*ls_args), particle_list=pparts,
                     initial_field=GaussianBeam(*beam_args, k_parallel_array=kpb, azimuthal_angles_array=al),
                     k_parallel=kp, solver_type='gmres', solver_tolerance=1e-10, store_coupling_matrix=False,
                     coupling_matrix_lookup_resolution=5,
                     coupling_matrix_interpolator_kind='linear', precision='double')
    sim.run()
    got = np.concatenate([p.scattered_field.coefficients for p in pparts])
    ls = ol.LayerSystem(*ls_args)
    lso = osys.LinearSystemOracle(plist, oin.GaussianBeam(*beam_args, kpb, al), ls, kp)
    table, rho = oc.radial_coupling_lookup_table(wl, plist, ls, kp, 5)
    W = oc.RadialLookupCPU(table, rho, plist, 'linear').dense()
    want = lso.solve_lu(W)
    assert np.abs(got - want).max() < 1e-6 * np.abs(want).max()
