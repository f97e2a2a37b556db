"""GPU parity of the explicit coupling matrix (CouplingMatrixExplicit,
linear_system.py:441-486) against the FP64 oracle blocks
(direct_coupling_block, direct_coupling.py:18-105, plus
layer_mediated_coupling_block, layer_mediated_coupling.py:215-340, with the
reference's quadrature), its dense matvecs and the solve that stands in for
the reference's LU path."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import layers as ol
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0
THICK, NIDX = [0, 250, 0], [1.5, 2 + 0.01j, 1]


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _contour():
    return ob.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 4],
                                                neff_resolution=2e-2)


def _particles():
    # three in the film, two above, one below; mixed multipole cut-offs. The
    # outer-layer particles sit within 80 nm of the film: a block between the
    # two outer layers carries the response entry of the wave that does not
    # exist there (zero up to roundoff in the reference's FP64 transfer
    # matrices) times e^{|Im kz| z} growth of evanescent kappa, so further out
    # both sides hold different amplified roundoff.
    spec = [([0, 0, 90], 3, 3), ([210, -60, 170], 3, 2), ([-150, 180, 60], 2, 2), ([60, 40, 330], 3, 3),
            ([-220, -90, 300], 2, 1), ([120, 200, -70], 3, 3)]
    return [osys.Sphere(p, 1.6 + 0.02j, 60, l, m) for p, l, m in spec]


def _operator(plist, thick, nidx, kp):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixExplicit
    return CouplingMatrixExplicit(WL, plist, LayerSystem(thick, nidx), k_parallel=kp)


def test_explicit_matrix_matches_oracle_across_layers(gpu):
    plist = _particles()
    kp = _contour()
    print('explicit matrix, 6 particles in 3 layers', flush=True)
    op = _operator(plist, THICK, NIDX, kp)
    want = osys.LinearSystemOracle(plist, osys.PlaneWave(WL, 0.3, 0.2, 0), ol.LayerSystem(THICK, NIDX),
                                   kp).explicit_coupling()
    got = op.coupling_matrix
    assert got.shape == want.shape
    err = np.abs(got - want).max() / np.abs(want).max()
    assert err < 5e-11, err


def test_explicit_matrix_one_medium_is_direct_only(gpu):
    # a layer system of one medium: direct coupling between every pair, no layer-mediated part
    plist = _particles()
    thick, nidx = [0, 0], [1.3, 1.3]
    print('explicit matrix, homogeneous medium', flush=True)
    op = _operator(plist, thick, nidx, _contour())
    want = osys.LinearSystemOracle(plist, osys.PlaneWave(WL, 0.3, 0.2, 0), ol.LayerSystem(thick, nidx),
                                   None).explicit_coupling()
    err = np.abs(op.coupling_matrix - want).max() / np.abs(want).max()
    assert err < 1e-12, err


def test_explicit_matvecs_match_dense_product(gpu):
    plist = _particles()
    op = _operator(plist, THICK, NIDX, _contour())
    W = op.coupling_matrix
    rng = np.random.default_rng(4)
    x = rng.standard_normal((3, W.shape[0])) + 1j * rng.standard_normal((3, W.shape[0]))
    want = x @ W.T
    scale = np.abs(want).max()
    assert np.abs(op.matvec(x, precision='double') - want).max() < 1e-13 * scale
    assert np.abs(op.matvec(x, precision='single') - want).max() < 2e-6 * scale


def test_explicit_lu_path_solves_the_system(gpu):
    # solver_type 'LU' (no lookup): GPU explicit matrix, FP64 GMRES to 1e-12, against the oracle's LU solve
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.simulation import Simulation
    plist = _particles()
    kp = _contour()
    pw = (WL, 0.4, 0.7, 1)
    ls_o = ol.LayerSystem(THICK, NIDX)
    lso = osys.LinearSystemOracle(plist, osys.PlaneWave(*pw), ls_o, kp)
    want = lso.solve_lu(lso.explicit_coupling())
    pparts = [Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in plist]
    sim = Simulation(layer_system=LayerSystem(THICK, NIDX), particle_list=pparts, initial_field=PlaneWave(*pw),
                     k_parallel=kp, solver_type='LU', store_coupling_matrix=True)
    sim.run()
    got = np.concatenate([p.scattered_field.coefficients for p in pparts])
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


def test_lu_solve_on_overlapping_spheres(gpu):
    """solver_type 'LU' is a direct solve (linear_system.py:259-268): on the overlapping-sphere system that
    stagnates restarted GMRES(20) near 3e-3, the device LU of M = 1 - T W (explicit matrix, one medium and a
    substrate) matches the oracle's dense solve of the same matrix to 1e-12, takes no Arnoldi steps, reuses
    its factors for a second call, and a lookup operator is refused as the reference refuses it."""
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixExplicit, CouplingMatrixRadialLookupCUDA
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    plist = [osys.Sphere([0, 0, 150], 2.4 + 0.1j, 100, 3, 3), osys.Sphere([102, -100, 150], 1.9 + 0.2j, 80, 3, 2),
             osys.Sphere([202, 100, 150], 1.7, 90, 3, 3)]
    ls = LayerSystem([0, 0], [1.52, 1])
    op = CouplingMatrixExplicit(WL, plist, ls, k_parallel=kp)
    W = op.coupling_matrix
    rng = np.random.default_rng(3)
    n = W.shape[0]
    t = 0.3 + 0.1j + 0.05 * rng.standard_normal(n)
    M = np.eye(n) - t[:, None] * W
    R = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    want = np.linalg.solve(M, R.T).T
    for call in range(2):
        X, st, ok = op.solve_master(t, R, tol=1e-4, precision='lu', with_stats=True)
        print('LU call %d:' % call, st, 'cond', np.linalg.cond(M), flush=True)
        assert ok and st['iterations'] == 0 and st['relative_residual'] < 1e-12
        for j in range(3):
            assert np.linalg.norm(X[j] - want[j]) / np.linalg.norm(want[j]) < 1e-12, (call, j)
    lookup = CouplingMatrixRadialLookupCUDA(WL, plist, ls, k_parallel=kp, resolution=5.0, interpolator_kind='cubic')
    from smuthi_amd import _lib
    with pytest.raises(_lib.SmhError):
        lookup.solve_master(t, R[0], tol=1e-4, precision='lu')
