"""GPU parity of the volume lookup (CouplingMatrixVolumeLookupCUDA,
linear_system.py:688-783; tables of prepare_lookup.py:16-250) against the FP64
oracle: the complex64 tables, the bilinear matvec and a solve."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import coupling as oc
from oracle import layers as ol
from oracle.system import Sphere as OSphere

pytestmark = pytest.mark.gpu

WL = 550.0
THICK, NIDX = [0, 400, 0], [1.5, 2 + 0.01j, 1]


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _film_particles(n, seed, l_max=2, m_max=2, side=900.0):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = np.array([rng.uniform(-side / 2, side / 2), rng.uniform(-side / 2, side / 2), rng.uniform(120, 280)])
        if all(np.linalg.norm(p - q) > 170 for q in pts):
            pts.append(p)
    return [OSphere(list(p), 1.6 + 0.02j, 60 + 10 * (i % 3), l_max, m_max) for i, p in enumerate(pts)]


def _contour():
    return ob.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 4],
                                                neff_resolution=2e-2)


def test_volume_tables_match_oracle(gpu):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_particles(5, 1)
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(THICK, NIDX), kp, 10.0)
    op = CouplingMatrixVolumeLookupCUDA(WL, plist, LayerSystem(THICK, NIDX), k_parallel=kp, resolution=10.0)
    np.testing.assert_allclose(op.rho_array, rho)
    np.testing.assert_allclose(op.sum_z_array, sz)
    np.testing.assert_allclose(op.diff_z_array, dz)
    for got, want in ((op.lookup_table_plus, tpl), (op.lookup_table_minus, tmn)):
        assert got.shape == want.shape
        scale = np.abs(want).max()
        # both are FP64 sums rounded to complex64: equal up to one float ulp (plus sum order)
        err = np.abs(got - want)
        assert (err <= 3e-7 * np.abs(want) + 3e-8 * scale).all(), err.max() / scale


def test_volume_matvec_matches_oracle(gpu):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_particles(12, 2)
    plist[3] = OSphere(plist[3].position, 1.7, 70, 2, 1)  # mixed multipole cut-offs
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(THICK, NIDX), kp, 10.0)
    ref = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, plist, 'linear')
    rng = np.random.default_rng(5)
    v = rng.standard_normal(ref.n_unk) + 1j * rng.standard_normal(ref.n_unk)
    want = ref.matvec(v)
    up = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0)
    built = CouplingMatrixVolumeLookupCUDA(WL, plist, LayerSystem(THICK, NIDX), k_parallel=kp, resolution=10.0)
    for op in (up, built):
        e64 = np.linalg.norm(op.matvec(v, 'double') - want) / np.linalg.norm(want)
        e32 = np.linalg.norm(op.matvec(v, 'single') - want) / np.linalg.norm(want)
        assert e64 < (1e-12 if op is up else 2e-6), e64
        assert e32 < 3e-6, e32


def test_volume_solve_matches_oracle_lu(gpu):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_particles(8, 3)
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(THICK, NIDX), kp, 10.0)
    W = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, plist, 'linear').dense()
    rng = np.random.default_rng(8)
    t = 0.2 + 0.1j + 0.05 * (rng.standard_normal(W.shape[0]) + 1j * rng.standard_normal(W.shape[0]))
    rhs = rng.standard_normal(W.shape[0]) + 1j * rng.standard_normal(W.shape[0])
    x_ref = np.linalg.solve(np.eye(W.shape[0]) - t[:, None] * W, rhs)
    op = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0)
    for mode in ('double', 'mixed'):
        x = op.solve_master(t, rhs, tol=1e-9, precision=mode)
        assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-7, mode


@pytest.mark.parametrize('cubic,okind', [('spline', 'cubic'), ('catmull-rom', 'catmull-rom')])
def test_volume_bicubic_matvec_matches_oracle(gpu, cubic, okind):
    # bicubic volume lookup: tensor not-a-knot spline (= RectBivariateSpline kx=ky=3,
    # linear_system.py:652-663) and tensor Catmull-Rom (linear_system_cuda.py:149-256)
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_particles(10, 4)
    plist[2] = OSphere(plist[2].position, 1.7, 70, 2, 1)  # mixed multipole cut-offs
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(THICK, NIDX), kp, 10.0)
    ref = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, plist, okind)
    rng = np.random.default_rng(9)
    v = rng.standard_normal(ref.n_unk) + 1j * rng.standard_normal(ref.n_unk)
    want = ref.matvec(v)
    up = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0, interpolator_kind='cubic',
                                                    cubic_interpolant=cubic)
    built = CouplingMatrixVolumeLookupCUDA(WL, plist, LayerSystem(THICK, NIDX), k_parallel=kp, resolution=10.0,
                                           interpolator_kind='cubic', cubic_interpolant=cubic)
    for op in (up, built):
        e64 = np.linalg.norm(op.matvec(v, 'double') - want) / np.linalg.norm(want)
        e32 = np.linalg.norm(op.matvec(v, 'single') - want) / np.linalg.norm(want)
        assert e64 < (1e-11 if op is up else 2e-6), e64
        assert e32 < 5e-6, e32
    # the mixed-precision solve on the bicubic operator against the oracle's LU
    W = ref.dense()
    t = 0.2 + 0.1j + 0.05 * (rng.standard_normal(W.shape[0]) + 1j * rng.standard_normal(W.shape[0]))
    x_ref = np.linalg.solve(np.eye(W.shape[0]) - t[:, None] * W, v)
    x = up.solve_master(t, v, tol=1e-9, precision='mixed')
    assert np.linalg.norm(x - x_ref) / np.linalg.norm(x_ref) < 1e-7
