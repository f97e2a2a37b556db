"""Batched volume lookup (kBatch right-hand sides per launch) and the batched
solve of several particle ensembles sharing one table, against the FP64
oracle: the same volume lookup operator (VolumeLookupCPU, linear_system.py:
593-685 / linear_system_cuda.py:26-256) assembled densely per ensemble and
solved by LU, with the right-hand sides of the oracle's plane waves."""
import ctypes

import numpy as np
import pytest

from oracle import basics as ob
from oracle import coupling as oc
from oracle import layers as ol
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0
FILM = ([0, 1000, 0], [1.52, 1.8, 1])  # a 1 um film (n = 1.8) on glass, as the batching configuration


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _contour():
    return ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=FILM[1], neff_resolution=2e-2)


def _film_ensemble(n, seed, l_max=3, density=2.78e-6):
    """Spheres r ~ U(80, 120) nm, n ~ U(1.5, 2.0) + 0.01i at z ~ U(150, 600) nm inside the film, in-plane
    density 2.78 per um^2, centres at least r_i + r_j + 20 nm apart."""
    rng = np.random.default_rng(seed)
    side = np.sqrt(n / density)
    r = rng.uniform(80, 120, n)
    pts = []
    while len(pts) < n:
        p = np.array([rng.uniform(-side / 2, side / 2), rng.uniform(-side / 2, side / 2), rng.uniform(150, 600)])
        k = len(pts)
        if all(np.linalg.norm(p - q) >= r[k] + r[j] + 20 for j, q in enumerate(pts)):
            pts.append(p)
    nre = rng.uniform(1.5, 2.0, n)
    return [osys.Sphere(list(p), complex(nre[i], 0.01), float(r[i]), l_max, l_max) for i, p in enumerate(pts)]


def _tables(op):
    from smuthi_amd import _lib
    from smuthi_amd import cuda
    shape = (len(op_grids(op)[0]), len(op_grids(op)[1]), op.operator.blocksize, op.operator.blocksize)
    plus = np.empty(shape, dtype=np.complex128)
    minus = np.empty(shape, dtype=np.complex128)
    _lib.check(_lib.load().smh_op_download_table(op.operator._op, _lib.dptr(plus), _lib.dptr(minus)), cuda.context(),
               'smh_op_download_table')
    return plus, minus


def op_grids(batch):
    return batch.rho_array, batch.sum_z_array, batch.diff_z_array


@pytest.mark.parametrize('kind,cubic,okind', [('linear', 'spline', 'linear'), ('cubic', 'catmull-rom', 'catmull-rom'),
                                              ('cubic', 'spline', 'cubic')])
def test_volume_batched_matvec_matches_oracle(gpu, kind, cubic, okind):
    """8-RHS volume matvec launches (a full batch and a partial one) in FP64 and FP32 against the oracle's
    dense volume operator on the same tables; mixed multipole cut-offs."""
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_ensemble(14, 11, l_max=2)
    plist[3] = osys.Sphere(plist[3].position, 1.7, 90, 2, 1)
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, plist, ol.LayerSystem(*FILM), kp, 10.0)
    W = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, plist, okind).dense()
    op = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0, interpolator_kind=kind,
                                                    cubic_interpolant=cubic)
    rng = np.random.default_rng(12)
    V = rng.standard_normal((11, W.shape[0])) + 1j * rng.standard_normal((11, W.shape[0]))
    V[9] *= 1e-3  # right-hand sides of different scale in one batch
    want = V @ W.T
    got64 = op.matvec(V, 'double')
    got32 = op.matvec(V, 'single')
    got16 = op.matvec(V, 'half')  # FP16 operands, FP32 accumulation
    worst16 = 0.0
    for r in range(11):
        e64 = np.linalg.norm(got64[r] - want[r]) / np.linalg.norm(want[r])
        e32 = np.linalg.norm(got32[r] - want[r]) / np.linalg.norm(want[r])
        e16 = np.linalg.norm(got16[r] - want[r]) / np.linalg.norm(want[r])
        worst16 = max(worst16, e16)
        assert e64 < 1e-11, (r, e64)
        assert e32 < 5e-6, (r, e32)
        assert e16 < 3e-3, (r, e16)
    print('%s/%s: worst FP16-operand error %.2e' % (kind, cubic, worst16), flush=True)


def test_volume_lmax4_batched_matvec_and_solve(gpu):
    """l_max = m_max = 4 (48-slot blocks: the FP32 volume kernel in 3 K-parts, FP64 in 6, FP16 in 3 parts of
    the result tiles) and one (3, 2) sphere: 8-RHS matvecs against the oracle's Catmull-Rom volume operator
    on the operator's own tables (built on the GPU), and a mixed16 solve against LU."""
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    ens = _film_ensemble(12, 17, l_max=4)
    ens[2] = osys.Sphere(ens[2].position, 1.7, 90, 3, 2)
    plist = [Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in ens]
    op = CouplingMatrixVolumeLookupCUDA(WL, plist, LayerSystem(*FILM), k_parallel=_contour(), resolution=10.0,
                                        interpolator_kind='cubic', cubic_interpolant='catmull-rom')
    assert op.blocksize == 48
    W = oc.VolumeLookupCPU(op.lookup_table_plus, op.lookup_table_minus, op.rho_array, op.sum_z_array,
                           op.diff_z_array, ens, 'catmull-rom').dense()
    rng = np.random.default_rng(21)
    V = rng.standard_normal((9, W.shape[0])) + 1j * rng.standard_normal((9, W.shape[0]))
    want = V @ W.T
    got = {p: op.matvec(V, p) for p in ('double', 'single', 'half')}
    one = op.matvec(V[4], 'double')
    assert np.linalg.norm(one - want[4]) / np.linalg.norm(want[4]) < 1e-11
    for r in range(9):
        nrm = np.linalg.norm(want[r])
        e = {p: np.linalg.norm(g[r] - want[r]) / nrm for p, g in got.items()}
        assert e['double'] < 1e-11 and e['single'] < 5e-6 and e['half'] < 3e-3, (r, e)
    t = np.concatenate([p.t_matrix(WL, FILM[1][1]).diagonal() for p in ens])
    M = np.eye(W.shape[0]) - t[:, None] * W
    X, st, ok = op.solve_master(t, V, tol=1e-9, precision='mixed16', with_stats=True)
    print('l_max 4 volume mixed16', st, flush=True)
    assert ok and st['matvecs_fp16'] > 0
    X_ref = np.linalg.solve(M, V.T).T
    for r in range(9):
        assert np.linalg.norm(X[r] - X_ref[r]) / np.linalg.norm(X_ref[r]) < 1e-7, r


def _fields(n_angles):
    from smuthi_amd.initial_field import PlaneWave
    th = np.arange(n_angles) * np.radians(60) / max(1, n_angles - 1)
    return [(WL, np.pi - t, 0.3, pol) for t in th for pol in (0, 1)], PlaneWave


@pytest.mark.parametrize('precision', ['double', 'mixed', 'mixed16'])
def test_ensembles_batched_solve_matches_oracle_lu(gpu, precision):
    """Four ensembles of 20-60 spheres inside a 1 um film share one volume table (Catmull-Rom, the
    reference's CUDA 'cubic'); 16 distinct plane waves (8 polar angles x TE/TM) are solved for every
    ensemble in one batched solve. Every (ensemble, field) system meets the FP64 tolerance and matches the
    oracle's LU of its own dense operator on the same table; the right-hand sides match the oracle's."""
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    kp = _contour()
    ens_o = [_film_ensemble(n, 40 + k) for k, n in enumerate((20, 35, 50, 60))]
    ens = [[Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in e] for e in ens_o]
    batch = EnsembleBatch(WL, ens, LayerSystem(*FILM), k_parallel=kp, resolution=10.0, interpolator_kind='cubic',
                          cubic_interpolant='catmull-rom')
    fargs, PlaneWave = _fields(8)
    rhs = batch.right_hand_sides([PlaneWave(*a) for a in fargs])
    x, st, ok = batch.solve_with_stats(rhs, tol=1e-10, precision=precision)
    print(precision, st, flush=True)
    assert ok and st['relative_residual'] <= 1e-10
    tpl, tmn = _tables(batch)
    ls_o = ol.LayerSystem(*FILM)
    for e, (plist, xe, be) in enumerate(zip(ens_o, batch.split(x), batch.split(rhs))):
        W = oc.VolumeLookupCPU(tpl, tmn, batch.rho_array, batch.sum_z_array, batch.diff_z_array, plist,
                               'catmull-rom').dense()
        lso = osys.LinearSystemOracle(plist, osys.PlaneWave(*fargs[0]), ls_o, kp)
        M = lso.master_dense(W)
        want = np.linalg.solve(M, be.T).T
        for j in range(len(fargs)):
            err = np.linalg.norm(xe[j] - want[j]) / np.linalg.norm(want[j])
            assert err < 1e-8, (e, j, err)
        # right-hand side of the first field against the oracle's T a
        r0 = osys.LinearSystemOracle(plist, osys.PlaneWave(*fargs[3]), ls_o, kp).right_hand_side()
        assert np.linalg.norm(be[3] - r0) / np.linalg.norm(r0) < 1e-10, e


def test_ensembles_equal_separate_operators(gpu):
    """An operator over three ensembles applies, ensemble by ensemble, what three separate operators on the
    same table apply (no coupling across ensembles), in single and batched launches; radial lookup."""
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    from smuthi_amd import ensembles
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    ens = [ensembles.monolayer(n, side, 200.0, seed=70 + n) for n, side in ((12, 2000.0), (30, 3500.0), (7, 1500.0))]
    batch = EnsembleBatch(WL, ens, LayerSystem([0, 0], [1.52, 1]), k_parallel=kp, resolution=5.0,
                          interpolator_kind='cubic', cubic_interpolant='spline')
    assert batch.kind == 'radial'
    from smuthi_amd import _lib
    B = batch.operator.blocksize
    table = np.empty((len(batch.radial_distance_array), B, B), dtype=np.complex128)
    _lib.check(_lib.load().smh_op_download_table(batch.operator._op, _lib.dptr(table), None), gpu.context(),
               'smh_op_download_table')
    rng = np.random.default_rng(2)
    n = batch.n_unknowns
    V = rng.standard_normal((9, n)) + 1j * rng.standard_normal((9, n))
    got = batch.operator.matvec(V, 'double')
    for e, (plist, ve, ge) in enumerate(zip(ens, batch.split(V), batch.split(got))):
        sep = CouplingMatrixRadialLookupCUDA.from_table(table, batch.radial_distance_array, plist, 5.0,
                                                        interpolator_kind='cubic')
        want = sep.matvec(ve, 'double')
        assert np.abs(ge - want).max() < 1e-12 * np.abs(want).max(), e
    got1 = batch.operator.matvec(V[0], 'double')
    assert np.abs(got1 - got[0]).max() < 1e-12 * np.abs(got[0]).max()


def test_preconditioned_solve_meets_the_tolerance(gpu):
    """A right block preconditioner changes the iteration, not the answer: exact block-Jacobi blocks
    (I - t_i W_ii)^-1 and a perturbed one both give solutions within 1e-8 of the unpreconditioned solve,
    certified by the FP64 residual; removing it restores the plain solve."""
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.particles import Sphere
    kp = _contour()
    ens_o = [_film_ensemble(n, 60 + n) for n in (25, 40)]
    ens = [[Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in e] for e in ens_o]
    batch = EnsembleBatch(WL, ens, LayerSystem(*FILM), k_parallel=kp, resolution=10.0, interpolator_kind='cubic',
                          cubic_interpolant='catmull-rom')
    fargs, PlaneWave = _fields(5)
    rhs = batch.right_hand_sides([PlaneWave(*a) for a in fargs])
    x0, st0, ok0 = batch.solve_with_stats(rhs, tol=1e-10, precision='mixed')
    assert ok0
    tpl, tmn = _tables(batch)
    B = batch.operator.blocksize
    blocks = np.zeros((len(batch.particle_list), B, B), complex)
    k = 0
    for plist in ens_o:
        ref = oc.VolumeLookupCPU(tpl, tmn, batch.rho_array, batch.sum_z_array, batch.diff_z_array, plist,
                                 'catmull-rom')
        for i, p in enumerate(plist):
            t = osys.Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max).t_matrix(WL, FILM[1][1])
            blocks[k] = np.linalg.inv(np.eye(B) - t @ ref.block(i, i))
            k += 1
    rng = np.random.default_rng(4)
    for P in (blocks, blocks * (1 + 0.05 * rng.standard_normal(blocks.shape))):
        batch.set_preconditioner(P)
        x, st, ok = batch.solve_with_stats(rhs, tol=1e-10, precision='mixed')
        print('preconditioned', st, 'plain', st0, flush=True)
        assert ok and st['relative_residual'] <= 1e-10
        assert np.linalg.norm(x - x0) / np.linalg.norm(x0) < 1e-8
    batch.set_preconditioner(None)
    x, st, ok = batch.solve_with_stats(rhs, tol=1e-10, precision='mixed')
    assert ok and np.linalg.norm(x - x0) / np.linalg.norm(x0) < 1e-9
