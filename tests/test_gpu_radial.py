"""GPU parity of the radial-lookup table builder and coupling matvec against
the FP64 oracle (restatement of prepare_lookup.radial_coupling_lookup_table and
CouplingMatrixRadialLookupCPU.matvec)."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import layers as ol
from oracle import coupling as oc
from oracle.system import Sphere as OSphere

pytestmark = pytest.mark.gpu

WL = 550


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _three_spheres():
    # geometry of tests/unit_tests/multiple_scattering_tests/test_radial_lookup.py:33-38
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 5],
                                              neff_resolution=1e-2)
    plist = [OSphere([0, 0, 150], 2.4 + 0.1j, 100, 3, 3), OSphere([102, -100, 150], 1.9 + 0.2j, 80, 3, 2),
             OSphere([202, 100, 150], 1.7, 90, 3, 3)]
    return kp, plist


def _random_spheres(n, seed, side=3000.0, z=200.0, mixed=True):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-side / 2, side / 2, 2)
        if all(np.hypot(*(p - q)) > 220 for q in pts):
            pts.append(p)
    orders = [(3, 3), (3, 2), (2, 2), (3, 1)] if mixed else [(3, 3)]
    return [OSphere([p[0], p[1], z], 1.5 + 0.01j, 100, *orders[i % len(orders)]) for i, p in enumerate(pts)]


def _product_op(plist, ls_args, kp, res, kind, cubic='spline'):
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    from smuthi_amd.layers import LayerSystem
    return CouplingMatrixRadialLookupCUDA(WL, plist, LayerSystem(*ls_args), k_parallel=kp, resolution=res,
                                          interpolator_kind=kind, cubic_interpolant=cubic)


LS3 = ([0, 400, 0], [1.5, 2 + 0.01j, 1])


def test_gpu_table_matches_oracle(gpu):
    kp, plist = _three_spheres()
    want, rho = oc.radial_coupling_lookup_table(WL, plist, ol.LayerSystem(*LS3), kp, 5.0)
    op = _product_op(plist, LS3, kp, 5.0, 'linear')
    np.testing.assert_array_equal(op.radial_distance_array, rho)
    got = op.lookup_table
    assert got.shape == want.shape
    scale = np.abs(want).max()
    # FP64 throughout; the direct part is rounded to complex64 (one float ulp may differ)
    err = np.abs(got - want)
    assert (err <= 2e-7 * np.abs(want) + 1e-12 * scale).all(), err.max() / scale


@pytest.mark.parametrize('order,res', [((1, 1), 5.0), ((2, 1), 5.0), ((2, 2), 5.0), ((3, 1), 5.0), ((3, 2), 5.0),
                                       ((4, 1), 5.0), ((4, 2), 2.0), ((4, 3), 2.0), ((4, 4), 2.0), ((4, 4), 5.0)])
def test_gpu_batched_matvec_orders(gpu, order, res):
    # every multipole cut-off of the 8-RHS slot-order kernels (l_max = 4 with m_max >= 2: 48-slot blocks,
    # the FP32 kernel in 3 K-parts, the FP64 kernel in 6), 9 right-hand sides = one full and one padded block
    rng = np.random.default_rng(31)
    pts = []
    while len(pts) < 16:
        p = rng.uniform(-1200, 1200, 2)
        if all(np.hypot(*(p - q)) > 220 for q in pts):
            pts.append(p)
    plist = [OSphere([p[0], p[1], 200.0], 1.5 + 0.01j, 100, *order) for p in pts]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = _product_op(plist, ([0, 0], [1.52, 1]), kp, res, 'cubic')
    ref_op = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic')
    V = rng.standard_normal((9, ref_op.n_unk)) + 1j * rng.standard_normal((9, ref_op.n_unk))
    W64 = op.matvec(V, precision='double')
    W32 = op.matvec(V, precision='single')
    # FP16 operands (FP16 interpolant coefficients scaled per bin, FP16 inputs scaled by max |x|; the
    # 48-slot blocks in 3 parts of the result tiles), linear or cubic inner operator
    W16 = op.matvec(V, precision='half')
    print(order, res, 'FP16 Horner terms', op.half_interpolation()[0], flush=True)
    for r in range(9):
        want = ref_op.matvec(V[r])
        assert np.linalg.norm(W64[r] - want) / np.linalg.norm(want) < 1e-12, r
        assert np.linalg.norm(W32[r] - want) / np.linalg.norm(want) < 3e-6, r
        assert np.linalg.norm(W16[r] - want) / np.linalg.norm(want) < 2e-3, r


@pytest.mark.parametrize('kind,cubic,okind', [('linear', 'spline', 'linear'), ('cubic', 'spline', 'cubic'),
                                              ('cubic', 'catmull-rom', 'catmull-rom')])
def test_gpu_matvec_matches_oracle_three_spheres(gpu, kind, cubic, okind):
    kp, plist = _three_spheres()
    op = _product_op(plist, LS3, kp, 5.0, kind, cubic)
    table = op.lookup_table
    ref_op = oc.RadialLookupCPU(table, op.radial_distance_array, plist, okind)
    rng = np.random.default_rng(11)
    v = rng.standard_normal(ref_op.n_unk) + 1j * rng.standard_normal(ref_op.n_unk)
    want = ref_op.matvec(v)
    got64 = op.matvec(v, precision='double')
    assert np.linalg.norm(got64 - want) / np.linalg.norm(want) < 1e-12
    got32 = op.matvec(v, precision='single')
    assert np.linalg.norm(got32 - want) / np.linalg.norm(want) < 2e-6
    # the test vector of test_radial_lookup.py:106
    v2 = np.arange(ref_op.n_unk).astype(complex)
    assert np.linalg.norm(op.matvec(v2, 'double') - ref_op.matvec(v2)) / np.linalg.norm(ref_op.matvec(v2)) < 1e-12


@pytest.mark.parametrize('kind,cubic,okind', [('linear', 'spline', 'linear'), ('cubic', 'spline', 'cubic'),
                                              ('cubic', 'catmull-rom', 'catmull-rom')])
def test_gpu_matvec_matches_oracle_random_ensemble(gpu, kind, cubic, okind):
    plist = _random_spheres(90, seed=4)
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    ls = ([0, 0], [1.52, 1])
    op = _product_op(plist, ls, kp, 5.0, kind, cubic)
    table = op.lookup_table
    want_table, rho = oc.radial_coupling_lookup_table(WL, plist, ol.LayerSystem(*ls), kp, 5.0)
    assert np.abs(table - want_table).max() <= 2e-7 * np.abs(want_table).max()
    ref_op = oc.RadialLookupCPU(table, rho, plist, okind)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(ref_op.n_unk) + 1j * rng.standard_normal(ref_op.n_unk)
    want = ref_op.matvec(v)
    got64 = op.matvec(v, precision='double')
    assert np.linalg.norm(got64 - want) / np.linalg.norm(want) < 1e-12
    got32 = op.matvec(v, precision='single')
    assert np.linalg.norm(got32 - want) / np.linalg.norm(want) < 2e-6
    # several right-hand sides in one call
    V = np.stack([v, 1j * v, v.conj()])
    W = op.matvec(V, precision='double')
    for r in range(3):
        assert np.linalg.norm(W[r] - ref_op.matvec(V[r])) / np.linalg.norm(want) < 1e-12


def test_gpu_matvec_two_particles_and_degenerate_stack(gpu):
    # minimal ensemble, homogeneous medium (no layer-mediated part)
    plist = [OSphere([0, 0, 0], 1.5, 100, 3, 3), OSphere([350.0, 123.0, 0], 1.6, 100, 3, 3)]
    ls = ([0, 0], [1.0, 1.0])
    op = _product_op(plist, ls, None, 2.0, 'cubic')
    want_table, rho = oc.radial_coupling_lookup_table(WL, plist, ol.LayerSystem(*ls), None, 2.0)
    got = op.lookup_table
    assert np.abs(got - want_table).max() <= 2e-7 * np.abs(want_table).max()
    ref_op = oc.RadialLookupCPU(got, rho, plist, 'cubic')
    v = np.linspace(-1, 1, ref_op.n_unk) + 0.5j
    assert np.linalg.norm(op.matvec(v, 'double') - ref_op.matvec(v)) / np.linalg.norm(ref_op.matvec(v)) < 1e-12
    # against the explicit direct coupling blocks: spline interpolation error only
    W = np.zeros((60, 60), complex)
    W[:30, 30:] = oc.direct_coupling_block(WL, plist[0], plist[1], ol.LayerSystem(*ls))
    W[30:, :30] = oc.direct_coupling_block(WL, plist[1], plist[0], ol.LayerSystem(*ls))
    assert np.linalg.norm(op.matvec(v, 'double') - W @ v) / np.linalg.norm(W @ v) < 5e-5


@pytest.mark.parametrize('kind,cubic,okind', [('linear', 'spline', 'linear'), ('cubic', 'spline', 'cubic'),
                                              ('cubic', 'catmull-rom', 'catmull-rom')])
def test_gpu_batched_matvec_matches_oracle(gpu, kind, cubic, okind):
    # 11 right-hand sides: one full block of 8 and one padded block of 3 through the batched kernel
    plist = _random_spheres(70, seed=6)
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = _product_op(plist, ([0, 0], [1.52, 1]), kp, 5.0, kind, cubic)
    ref_op = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, okind)
    rng = np.random.default_rng(21)
    V = rng.standard_normal((11, ref_op.n_unk)) + 1j * rng.standard_normal((11, ref_op.n_unk))
    V[4] *= 1e-3  # right-hand sides of different scale share one fixed-point scale per block
    W64 = op.matvec(V, precision='double')
    W32 = op.matvec(V, precision='single')
    for r in range(11):
        want = ref_op.matvec(V[r])
        assert np.linalg.norm(W64[r] - want) / np.linalg.norm(want) < 1e-12, r
        assert np.linalg.norm(W32[r] - want) / np.linalg.norm(want) < 3e-6, r
