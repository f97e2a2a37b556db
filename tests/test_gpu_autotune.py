"""Launch-shape autotuner (smh_op_autotune): every candidate grouping of the
pair bins into work items gives the same matvec (within each kernel's
accuracy against the oracle, CouplingMatrixRadialLookupCPU.matvec,
linear_system.py:903-975, and the volume lookup, linear_system.py:593-783),
the tuner keeps the fastest one, and a solve after it (its captured graphs
dropped and rebuilt) still matches the oracle's LU."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import coupling as oc
from oracle import layers as ol
from oracle.system import Sphere as OSphere

from test_gpu_radial import WL, _product_op, _random_spheres

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


# (items per CU, launch order): 1 XCD-dealt, 0 chunk order, n >= 2 row bands (4096: about one
# band per table row, every entry its own item)
CANDIDATES = [(2, 1), (16, 1), (16, 0), (64, 1), (16, 8), (4, 64), (2, 4096)]


def _check_tune(res):
    assert set(res['matvec_ms']) == set(CANDIDATES)
    assert all(v > 0 for v in res['matvec_ms'].values())
    assert res['matvec_ms'][res['best']] == min(res['matvec_ms'].values())


def test_radial_autotune_keeps_results(gpu):
    plist = _random_spheres(120, seed=9, mixed=False)
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    op = _product_op(plist, ([0, 0], [1.52, 1]), kp, 5.0, 'cubic', 'spline')
    ref = oc.RadialLookupCPU(op.lookup_table, op.radial_distance_array, plist, 'cubic')
    rng = np.random.default_rng(3)
    V = rng.standard_normal((8, ref.n_unk)) + 1j * rng.standard_normal((8, ref.n_unk))
    want = np.stack([ref.matvec(v) for v in V])
    scale = np.linalg.norm(want, axis=1)

    def errors():
        e64 = np.linalg.norm(op.matvec(V[0], 'double') - want[0]) / scale[0]
        e32 = np.linalg.norm(op.matvec(V[0], 'single') - want[0]) / scale[0]
        b64 = np.max(np.linalg.norm(op.matvec(V, 'double') - want, axis=1) / scale)
        b32 = np.max(np.linalg.norm(op.matvec(V, 'single') - want, axis=1) / scale)
        b16 = np.max(np.linalg.norm(op.matvec(V, 'half') - want, axis=1) / scale)
        return e64, e32, b64, b32, b16

    for shape in ((False, 'double'), (False, 'single'), (True, 'half'), (True, 'double')):
        res = op.autotune(batch=shape[0], precision=shape[1], iters=2, candidates=CANDIDATES)
        print('autotune', shape, {k: round(v, 4) for k, v in res['matvec_ms'].items()}, 'best', res['best'],
              flush=True)
        _check_tune(res)
        e64, e32, b64, b32, b16 = errors()
        print('  errors vs oracle: %.1e %.1e %.1e %.1e %.1e' % (e64, e32, b64, b32, b16), flush=True)
        assert e64 < 1e-12 and b64 < 1e-12
        assert e32 < 2e-6 and b32 < 3e-6
        assert b16 < 2e-3
    # every candidate in turn for the FP64 single-vector kernel: results stay put
    for c in CANDIDATES:
        op.autotune(batch=False, precision='double', iters=1, candidates=[c])
        assert np.linalg.norm(op.matvec(V[1], 'double') - want[1]) / scale[1] < 1e-12, c
    # launch orders outside [0, 4096] and item counts outside [1, 1024] are rejected
    from smuthi_amd._lib import SmhError
    for bad in ((16, 4097), (16, -1), (0, 1), (1025, 1)):
        with pytest.raises(SmhError):
            op.autotune(batch=True, precision='double', iters=1, candidates=[bad])
    assert np.linalg.norm(op.matvec(V[1], 'double') - want[1]) / scale[1] < 1e-12
    # the solver's graphs were dropped with the old item arrays: a solve still matches LU
    t = 0.05 + 0.02j + 0.01 * rng.standard_normal(ref.n_unk)  # a well-conditioned M (GMRES(20) converges)
    # the dense W from FP64 matvecs of the unit vectors (checked against the oracle above)
    W = op.matvec(np.eye(ref.n_unk, dtype=complex), 'double').T
    M = np.eye(ref.n_unk) - t[:, None] * W
    R = V[:3]
    X = op.solve_master(t, R, tol=1e-10, precision='mixed')
    op.autotune(batch=True, precision='single', iters=1, candidates=[(4, 0)])
    X2 = op.solve_master(t, R, tol=1e-10, precision='mixed')
    for j in range(3):
        x_lu = np.linalg.solve(M, R[j])
        assert np.linalg.norm(X[j] - x_lu) / np.linalg.norm(x_lu) < 1e-8
        assert np.linalg.norm(X2[j] - x_lu) / np.linalg.norm(x_lu) < 1e-8


def test_volume_autotune_keeps_results(gpu):
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    rng = np.random.default_rng(5)
    pts = []
    while len(pts) < 40:
        p = np.array([rng.uniform(-600, 600), rng.uniform(-600, 600), rng.uniform(120, 280)])
        if all(np.linalg.norm(p - q) > 170 for q in pts):
            pts.append(p)
    plist = [OSphere(list(p), 1.6 + 0.02j, 70, 2, 2) for p in pts]
    thick, nidx = [0, 400, 0], [1.5, 2 + 0.01j, 1]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, neff_waypoints=[0, 0.5, 0.8 - 0.01j, 2 - 0.01j, 2.5, 4],
                                              neff_resolution=2e-2)
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(thick, nidx), kp, 10.0)
    ref = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, plist, 'catmull-rom')
    op = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0, interpolator_kind='cubic',
                                                    cubic_interpolant='catmull-rom')
    V = rng.standard_normal((8, ref.n_unk)) + 1j * rng.standard_normal((8, ref.n_unk))
    want = np.stack([ref.matvec(v) for v in V])
    scale = np.linalg.norm(want, axis=1)
    for shape in ((True, 'single'), (True, 'double'), (False, 'single'), (True, 'half')):
        res = op.autotune(batch=shape[0], precision=shape[1], iters=2, candidates=CANDIDATES)
        print('volume autotune', shape, res['best'], flush=True)
        _check_tune(res)
        b64 = np.max(np.linalg.norm(op.matvec(V, 'double') - want, axis=1) / scale)
        b32 = np.max(np.linalg.norm(op.matvec(V, 'single') - want, axis=1) / scale)
        b16 = np.max(np.linalg.norm(op.matvec(V, 'half') - want, axis=1) / scale)
        e32 = np.linalg.norm(op.matvec(V[0], 'single') - want[0]) / scale[0]
        assert b64 < 1e-12 and b32 < 3e-6 and e32 < 3e-6 and b16 < 3e-3, (b64, b32, e32, b16)
    # every candidate in turn for the FP64 8-RHS kernel: results stay put
    for c in CANDIDATES:
        op.autotune(batch=True, precision='double', iters=1, candidates=[c])
        b64 = np.max(np.linalg.norm(op.matvec(V, 'double') - want, axis=1) / scale)
        assert b64 < 1e-12, c
