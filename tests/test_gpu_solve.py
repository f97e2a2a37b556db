"""GPU GMRES solves (plain FP64 and FP32-in-FP64 refinement) against the
FP64 oracle: the same lookup operator assembled densely and solved by LU."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import layers as ol
from oracle import coupling as oc
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _run_product(plist_o, ls_args, kp, pw_args, res, kind, cubic, tol, precision):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.simulation import Simulation
    pparts = [Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in plist_o]
    sim = Simulation(layer_system=LayerSystem(*ls_args), particle_list=pparts, initial_field=PlaneWave(*pw_args),
                     k_parallel=kp, solver_type='gmres', solver_tolerance=tol, store_coupling_matrix=False,
                     coupling_matrix_lookup_resolution=res, coupling_matrix_interpolator_kind=kind,
                     cubic_interpolant=cubic, precision=precision)
    sim.run()
    b = np.concatenate([p.scattered_field.coefficients for p in pparts])
    return sim, b


def _oracle_lookup_solution(plist_o, ls_args, kp, pw_args, res, okind):
    ls = ol.LayerSystem(*ls_args)
    table, rho = oc.radial_coupling_lookup_table(WL, plist_o, ls, kp, res)
    W = oc.RadialLookupCPU(table, rho, plist_o, okind).dense()
    lsys = osys.LinearSystemOracle(plist_o, osys.PlaneWave(*pw_args), ls, kp)
    return lsys, W, lsys.solve_lu(W)


@pytest.mark.parametrize('precision', ['double', 'mixed'])
@pytest.mark.parametrize('kind,cubic,okind', [('linear', 'spline', 'linear'), ('cubic', 'spline', 'cubic'),
                                              ('cubic', 'catmull-rom', 'catmull-rom')])
def test_three_spheres_solution_matches_oracle(gpu, precision, kind, cubic, okind):
    # geometry and tolerances of tests/unit_tests/multiple_scattering_tests/test_radial_lookup.py
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 5],
                                              neff_resolution=1e-2)
    plist = [osys.Sphere([0, 0, 150], 2.4 + 0.1j, 100, 3, 3), osys.Sphere([102, -100, 150], 1.9 + 0.2j, 80, 3, 2),
             osys.Sphere([202, 100, 150], 1.7, 90, 3, 3)]
    ls_args = ([0, 400, 0], [1.5, 2 + 0.01j, 1])
    pw_args = (WL, np.pi * 7 / 8, np.pi / 3, 0)
    sim, b = _run_product(plist, ls_args, kp, pw_args, 5.0, kind, cubic, 1e-8, precision)
    lsys, W, b_ref = _oracle_lookup_solution(plist, ls_args, kp, pw_args, 5.0, okind)
    assert np.linalg.norm(b - b_ref) / np.linalg.norm(b_ref) < 1e-6
    st = sim.linear_system.solve_stats
    assert st['relative_residual'] <= 1e-8
    # the lookup solution against the explicit-coupling solution (test_radial_lookup.py:127-140)
    b_exp = lsys.solve_lu(lsys.explicit_coupling())
    n0 = 30
    assert np.linalg.norm(b[:n0] - b_exp[:n0]) / np.linalg.norm(b_exp[:n0]) < 5e-4


@pytest.mark.parametrize('precision', ['double', 'mixed'])
def test_fifty_spheres_on_substrate(gpu, precision):
    """The CPU-runnable configuration: 50 spheres on glass, radial linear 5 nm, tol 1e-5."""
    from smuthi_amd import ensembles
    pp = ensembles.small_substrate_ensemble(seed=1)
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 3, 3) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    ls_args = ([0, 0], [1.52, 1])
    pw_args = (WL, np.pi, 0, 0)
    sim, b = _run_product(plist, ls_args, kp, pw_args, 5.0, 'linear', 'spline', 1e-7, precision)
    lsys, W, b_ref = _oracle_lookup_solution(plist, ls_args, kp, pw_args, 5.0, 'linear')
    err = np.linalg.norm(b - b_ref) / np.linalg.norm(b_ref)
    assert err < 1e-5, err
    rhs = lsys.right_hand_side()
    M = lsys.master_dense(W)
    assert np.linalg.norm(rhs - M @ b) / np.linalg.norm(rhs) < 1e-6
    # scipy GMRES at smuthi's settings reaches the same coefficients within its tolerance
    b_gm, info = lsys.solve_gmres(lambda v: W @ v, tol=1e-7)
    assert np.linalg.norm(b - b_gm) / np.linalg.norm(b_gm) < 1e-5


def test_multiple_rhs_match_single(gpu):
    """Several incidence angles in one smh_solve call equal one-by-one solves."""
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.linear_system import LinearSystem
    from smuthi_amd import ensembles
    pparts = ensembles.small_substrate_ensemble(seed=3)[:20]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    ls = LayerSystem([0, 0], [1.52, 1])
    rhs = []
    sols = []
    for j, ang in enumerate((np.pi, np.pi * 0.9, np.pi * 0.8)):
        lsys = LinearSystem(pparts, PlaneWave(WL, ang, 0.3 * j, j % 2), ls, kp, solver_type='gmres',
                            solver_tolerance=1e-8, store_coupling_matrix=False, coupling_matrix_lookup_resolution=5.0,
                            interpolator_kind='cubic')
        lsys.prepare()
        rhs.append(lsys.t_matrix.right_hand_side())
        sols.append(lsys.solve())
    X = lsys.solve(np.stack(rhs))
    for j in range(3):
        assert np.linalg.norm(X[j] - sols[j]) / np.linalg.norm(sols[j]) < 1e-7


@pytest.mark.parametrize('precision', ['double', 'mixed', 'mixed16'])
def test_batched_solve_matches_oracle_lu(gpu, precision):
    """Ten right-hand sides (one full batch of 8 and one of 2) against the oracle's LU."""
    from smuthi_amd import ensembles
    pp = ensembles.small_substrate_ensemble(seed=2)[:30]
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 3, 3) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=5.0,
                                        interpolator_kind='cubic')
    W = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic').dense()
    rng = np.random.default_rng(4)
    n = W.shape[0]
    t = 0.3 + 0.1j + 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    R = rng.standard_normal((10, n)) + 1j * rng.standard_normal((10, n))
    R[7] *= 1e3  # right-hand sides of different scale in one batch
    M = np.eye(n) - t[:, None] * W
    X_ref = np.linalg.solve(M, R.T).T
    X, stats, ok = op.solve_master(t, R, tol=1e-9, precision=precision, with_stats=True)
    assert ok and stats['relative_residual'] <= 1e-9
    # a second call replays the captured Arnoldi-step graphs: same steps, same matvec counts
    X2, stats2, ok2 = op.solve_master(t, R, tol=1e-9, precision=precision, with_stats=True)
    assert ok2 and np.abs(X2 - X).max() <= 1e-12 * np.abs(X).max()
    for key in ('iterations', 'matvecs_fp64', 'matvecs_fp32', 'matvecs_fp16', 'sweeps'):
        assert stats2[key] == stats[key], (key, stats, stats2)
    if precision == 'mixed16':
        # FP16 inner matvecs; FP32 residuals between sweeps; FP64 certificates
        assert stats['matvecs_fp16'] > 0 and stats['matvecs_fp64'] >= 1
    for j in range(10):
        assert np.linalg.norm(X[j] - X_ref[j]) / np.linalg.norm(X_ref[j]) < 1e-7, j
        assert np.linalg.norm(R[j] - M @ X[j]) / np.linalg.norm(R[j]) < 2e-9, j



@pytest.mark.parametrize('precision', ['mixed16', 'double'])
def test_device_resident_solve_matches_host(gpu, precision):
    """Right-hand sides, T and the solution in HBM (cuda.DeviceArray; smh_solve copies device to device): the
    same solutions as from host arrays, against the oracle's LU; a single device row too."""
    from smuthi_amd import ensembles
    from smuthi_amd.cuda import DeviceArray
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    pp = ensembles.small_substrate_ensemble(seed=3)[:24]
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 3, 3) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=5.0,
                                        interpolator_kind='cubic')
    W = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic').dense()
    rng = np.random.default_rng(9)
    n = W.shape[0]
    t = 0.3 + 0.1j + 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    R = rng.standard_normal((9, n)) + 1j * rng.standard_normal((9, n))
    X_ref = np.linalg.solve(np.eye(n) - t[:, None] * W, R.T).T
    X_host = op.solve_master(t, R, tol=1e-9, precision=precision)
    R_dev = DeviceArray.from_host(R)
    t_dev = DeviceArray.from_host(t)
    np.testing.assert_array_equal(R_dev.to_host(), R)
    np.testing.assert_array_equal(R_dev[2:5].to_host(), R[2:5])
    X_dev, st, ok = op.solve_master(t_dev, R_dev, tol=1e-9, precision=precision, with_stats=True)
    assert ok and isinstance(X_dev, DeviceArray) and X_dev.shape == R.shape
    X_dev = X_dev.to_host()
    x1 = op.solve_master(t_dev, R_dev[4], tol=1e-9, precision=precision).to_host()
    for j in range(9):
        assert np.linalg.norm(X_dev[j] - X_ref[j]) / np.linalg.norm(X_ref[j]) < 1e-7, j
        assert np.linalg.norm(X_dev[j] - X_host[j]) / np.linalg.norm(X_ref[j]) < 1e-7, j
    assert np.linalg.norm(x1 - X_ref[4]) / np.linalg.norm(X_ref[4]) < 1e-7


@pytest.mark.parametrize('precision,res', [('double', 5.0), ('mixed', 5.0), ('mixed16', 2.0), ('mixed16', 5.0)])
def test_lmax4_batched_solve_matches_oracle_lu(gpu, precision, res):
    """l_max = m_max = 4 (48 coefficients per particle, the slot-order 8-RHS kernels in K-parts, the FP16
    kernel in parts of the result tiles): ten right-hand sides against the oracle's LU of the same lookup
    operator; single-RHS solve too."""
    from smuthi_amd import ensembles
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    pp = ensembles.small_substrate_ensemble(seed=4)[:20]
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 4, 4) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=res,
                                        interpolator_kind='cubic')
    assert op.blocksize == 48
    W = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic').dense()
    rng = np.random.default_rng(14)
    n = W.shape[0]
    t = np.concatenate([p.t_matrix(WL, 1.0).diagonal() for p in plist])
    R = rng.standard_normal((10, n)) + 1j * rng.standard_normal((10, n))
    M = np.eye(n) - t[:, None] * W
    X_ref = np.linalg.solve(M, R.T).T
    X, st, ok = op.solve_master(t, R, tol=1e-9, precision=precision, with_stats=True)
    print(precision, st, flush=True)
    assert ok and st['relative_residual'] <= 1e-9
    if precision == 'mixed16':
        assert st['matvecs_fp16'] > 0
    x1 = op.solve_master(t, R[3], tol=1e-9, precision=precision)
    for j in range(10):
        assert np.linalg.norm(X[j] - X_ref[j]) / np.linalg.norm(X_ref[j]) < 1e-7, j
    assert np.linalg.norm(x1 - X_ref[3]) / np.linalg.norm(X_ref[3]) < 1e-7

def test_single_rhs_mixed16_solve(gpu):
    """One right-hand side with FP16-operand inner matvecs (padded into one batched launch)."""
    from smuthi_amd import ensembles
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    pp = ensembles.small_substrate_ensemble(seed=3)[:20]
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 3, 2) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=5.0,
                                        interpolator_kind='linear')
    W = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'linear').dense()
    rng = np.random.default_rng(6)
    n = W.shape[0]
    t = 0.25 - 0.1j + 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    M = np.eye(n) - t[:, None] * W
    x, stats, ok = op.solve_master(t, r, tol=1e-6, precision='mixed16', with_stats=True)
    assert ok and stats['relative_residual'] <= 1e-6 and stats['matvecs_fp16'] > 0
    assert np.linalg.norm(x - np.linalg.solve(M, r)) / np.linalg.norm(x) < 1e-5


def test_mixed16_linear_fp16_operator(gpu):
    """At a 2 nm table the FP16 inner operator interpolates linearly between the nodes (the error, measured on
    the table, stays within 2^-10 of every used bin's scale): its matvec stays within the FP16 tolerance of the
    cubic oracle, and mixed16 solves still certify against the cubic FP64 operator."""
    from smuthi_amd import ensembles
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    pp = ensembles.small_substrate_ensemble(seed=2)[:24]
    plist = [osys.Sphere(p.position, p.refractive_index, p.radius, 3, 3) for p in pp]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=2.0,
                                        interpolator_kind='cubic')
    terms, ratio = op.half_interpolation()
    print('FP16 operator: %d terms, linear error / bin scale %.3e' % (terms, ratio), flush=True)
    assert terms == 2 and 0.0 <= ratio <= 2.0 ** -10
    ref = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic')
    rng = np.random.default_rng(8)
    n = ref.n_unk
    V = rng.standard_normal((9, n)) + 1j * rng.standard_normal((9, n))
    W16 = op.matvec(V, precision='half')
    for r in range(9):
        want = ref.matvec(V[r])
        assert np.linalg.norm(W16[r] - want) / np.linalg.norm(want) < 2e-3, r
    W = ref.dense()
    t = 0.3 + 0.1j + 0.05 * (rng.standard_normal(n) + 1j * rng.standard_normal(n))
    R = rng.standard_normal((8, n)) + 1j * rng.standard_normal((8, n))
    M = np.eye(n) - t[:, None] * W
    X, stats, ok = op.solve_master(t, R, tol=1e-9, precision='mixed16', with_stats=True)
    assert ok and stats['relative_residual'] <= 1e-9 and stats['matvecs_fp16'] > 0
    X_ref = np.linalg.solve(M, R.T).T
    for j in range(8):
        assert np.linalg.norm(X[j] - X_ref[j]) / np.linalg.norm(X_ref[j]) < 1e-7, j


def test_refinement_on_ill_scaled_operator(gpu):
    """Overlapping spheres (centres 143 nm apart, radii 100 and 80 nm) make the lookup operator ill-scaled
    (entries up to ~340, cond(M) ~ 500): with restarts longer than the system (GMRES(20) stagnates near 3e-3
    on it, as scipy's would), the refinement (one inner restart cycle per sweep, FP64 residuals in between)
    meets the tolerance against the oracle's dense LU in FP32 and FP16 inner precision (single and batched)."""
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    plist = [osys.Sphere([0, 0, 150], 2.4 + 0.1j, 100, 3, 3), osys.Sphere([102, -100, 150], 1.9 + 0.2j, 80, 3, 2),
             osys.Sphere([202, 100, 150], 1.7, 90, 3, 3)]
    op = CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=5.0,
                                        interpolator_kind='cubic')
    ref = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic')
    rng = np.random.default_rng(0)
    n = ref.n_unk
    t = 0.3 + 0.1j + 0.05 * rng.standard_normal(n)
    M = np.eye(n) - t[:, None] * ref.dense()
    r = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, st, ok = op.solve_master(t, r, tol=1e-8, precision='mixed', restart=100, with_stats=True)
    print('single mixed:', st, flush=True)
    assert ok and st['relative_residual'] <= 1e-8
    assert np.linalg.norm(x - np.linalg.solve(M, r)) / np.linalg.norm(x) < 1e-6
    R = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    X_ref = np.linalg.solve(M, R.T).T
    for prec in ('mixed', 'mixed16'):
        X, st, ok = op.solve_master(t, R, tol=1e-8, precision=prec, restart=100, with_stats=True)
        print('batched %s:' % prec, st, flush=True)
        assert ok and st['relative_residual'] <= 1e-8, prec
        for j in range(3):
            assert np.linalg.norm(X[j] - X_ref[j]) / np.linalg.norm(X_ref[j]) < 1e-6, (prec, j)

