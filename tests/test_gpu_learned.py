"""Learned near-field preconditioner on the GPU against the CPU oracle
(oracle/mlp.py, pinned by finite differences; oracle/coupling.py lookup
operators): the network's forward pass and gradients (BF16 MFMA kernels),
the near-field blocks, the training loss and its gradient, and the solve:
preconditioned solves give the plain solve's answer under the same FP64
certificate, and the trained network cuts GMRES iterations on held-out
ensembles against plain GMRES."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import coupling as oc
from oracle import layers as ol
from oracle import mlp as om
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0
STACK = ([0, 500, 0], [1.52, 2.0, 1])  # air / 500 nm film (n = 2.0) / glass


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _contour():
    return ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=STACK[1], neff_resolution=2e-2)


def test_network_forward_and_gradients_match_oracle(gpu):
    """Production shape (K = 16: 136 inputs, hidden 256, 204 outputs), random parameters: outputs and the
    parameter gradient of sum(dy . y) within 1e-4 of the oracle's (same BF16 roundings, FP64 sums)."""
    from smuthi_amd.learned import LearnedPreconditioner
    lp = LearnedPreconditioner(K=16, hidden=256, l_max=3, seed=1)
    rng = np.random.default_rng(5)
    p = (rng.standard_normal(lp.n_params) * 0.05).astype(np.float32)
    lp.set_params(p)
    x = rng.standard_normal((300, lp.n_in)).astype(np.float32)
    dy = rng.standard_normal((300, lp.n_out)).astype(np.float32)
    y = lp.forward(x)
    y_ref = om.forward(p, x, lp.n_in, lp.hidden, lp.n_out)[0]
    ey = np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref)
    g = lp.backward(x, dy)
    g_ref = om.backward(p, x, dy, lp.n_in, lp.hidden, lp.n_out)
    eg = np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    print('network vs oracle: outputs %.2e, gradient %.2e' % (ey, eg), flush=True)
    assert ey < 1e-4 and eg < 1e-4
    # every parameter block individually
    sizes = [lp.n_in * 256, 256, 256 * 256, 256, 256 * lp.n_out, lp.n_out]
    o = 0
    for k, n in enumerate(sizes):
        e = np.linalg.norm(g[o:o + n] - g_ref[o:o + n]) / max(np.linalg.norm(g_ref[o:o + n]), 1e-30)
        assert e < 1e-4, (k, e)
        o += n


def _film_particles(n, seed):
    from smuthi_amd import ensembles
    return ensembles.learned_ensemble(n, seed)


def _as_oracle(plist):
    return [osys.Sphere(p.position, p.refractive_index, p.radius, p.l_max, p.m_max) for p in plist]


def test_near_blocks_and_loss_match_oracle(gpu):
    """Near-field blocks diag(t_i) W_ij of a volume operator (Catmull-Rom) against the oracle's lookup blocks,
    and the training loss and dL/dy on them against the oracle's."""
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.learned import LearnedPreconditioner
    from smuthi_amd.volume import CouplingMatrixVolumeLookupCUDA
    plist = _film_particles(40, 3)
    po = _as_oracle(plist)
    kp = _contour()
    tpl, tmn, rho, sz, dz = oc.volumetric_coupling_lookup_table(WL, po, ol.LayerSystem(*STACK), kp, 10.0)
    ref = oc.VolumeLookupCPU(tpl, tmn, rho, sz, dz, po, 'catmull-rom')
    op = CouplingMatrixVolumeLookupCUDA.from_tables(tpl, tmn, rho, sz, dz, plist, 10.0, interpolator_kind='cubic',
                                                    cubic_interpolant='catmull-rom')
    tb = [p.t_matrix(WL, STACK[1][1]) for p in po]
    t = np.concatenate([np.diag(b) for b in tb])
    lp = LearnedPreconditioner(K=16, hidden=256, l_max=3, seed=2)
    pos = np.array([p.position for p in plist])
    nbr, feat = lp.prepare(op, pos, [p.radius for p in plist], [p.refractive_index for p in plist], 0.0, 500.0, WL, t)
    from smuthi_amd import _lib
    B = op.blocksize
    A = np.empty((len(plist), 17, B, B), dtype=np.complex128)
    _lib.check(_lib.load().smh_op_download_near_blocks(op._op, _lib.dptr(A)), gpu.context(), 'download')
    scale = 0.0
    err = 0.0
    for i in range(len(plist)):
        for k in range(17):
            want = tb[i] @ ref.block(i, nbr[i, k])
            scale = max(scale, np.abs(want).max())
            err = max(err, np.abs(A[i, k] - want).max())
    print('near blocks: max error %.2e of %.2e' % (err, scale), flush=True)
    assert err < 1e-10 * scale
    # loss: random right-hand sides, arbitrary targets
    rng = np.random.default_rng(8)
    n = op.shape[0]
    R = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    X = rng.standard_normal((4, n)) + 1j * rng.standard_normal((4, n))
    lp.add_training_ensemble(op, feat, R, X)
    y = (0.1 * rng.standard_normal((len(plist), lp.n_out))).astype(np.float32)
    ch = om.channels(3, 3)
    U = om.near_products(A, nbr, R.reshape(4, len(plist), B))
    E0 = (R - X).reshape(4, len(plist), B).transpose(1, 0, 2)
    for w in (0.0, 0.05):
        lp.set_prior_weight(w)
        L, dy = lp.data_loss(y)
        L_ref, dy_ref = om.loss(y, U, E0, ch, float(np.sum(np.abs(X) ** 2)), w)
        print('loss (prior weight %g) %.6e vs %.6e' % (w, L, L_ref), flush=True)
        assert abs(L - L_ref) < 1e-5 * L_ref
        assert np.linalg.norm(dy - dy_ref) < 1e-4 * np.linalg.norm(dy_ref)


def _training_batch(ens_sizes, seeds, kp):
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.layers import LayerSystem
    ens = [_film_particles(n, s) for n, s in zip(ens_sizes, seeds)]
    return EnsembleBatch(WL, ens, LayerSystem(*STACK), k_parallel=kp, resolution=10.0, interpolator_kind='cubic',
                         cubic_interpolant='catmull-rom')


def _prepare(lp, batch):
    parts = batch.particle_list
    return lp.prepare(batch.operator, np.array([p.position for p in parts]), [p.radius for p in parts],
                      [p.refractive_index for p in parts], 0.0, 500.0, WL, batch.t_diagonal,
                      ensemble_offsets=batch.particle_offsets)


def _plane_waves(n_angles):
    from smuthi_amd.initial_field import PlaneWave
    th = np.radians(np.linspace(0, 60, n_angles))
    return [PlaneWave(WL, np.pi - t, 0.2, pol) for t in th for pol in (0, 1)]


def test_learned_preconditioner_cuts_iterations(gpu):
    """Train on solutions of 4 ensembles (random and plane-wave right-hand sides), evaluate on 2 held-out
    ensembles with 8 plane waves: the preconditioned solves (untrained structure and trained network)
    return the plain solve's answer under the FP64 certificate; training lowers the fit loss, and the
    trained network takes fewer GMRES iterations than plain GMRES."""
    from smuthi_amd.learned import LearnedPreconditioner, solve_with_fallback
    kp = _contour()
    rng = np.random.default_rng(11)
    lp = LearnedPreconditioner(K=16, hidden=256, l_max=3, seed=3)
    train = _training_batch([150, 180, 120, 160], [101, 102, 103, 104], kp)
    nbr, feat = _prepare(lp, train)
    n = train.n_unknowns
    R = np.concatenate([rng.standard_normal((8, n)) + 1j * rng.standard_normal((8, n)),
                        train.right_hand_sides(_plane_waves(4))])
    X, st, ok = train.solve_with_stats(R, tol=1e-9, precision='mixed')
    assert ok
    lp.add_training_ensemble(train.operator, feat, R, X)
    L0 = lp.data_loss()[0]
    losses = lp.train(300, lr=1e-3)
    L1 = lp.data_loss()[0]
    print('fit loss: untrained %.4e, trained %.4e (first/last step %.4e / %.4e)' % (L0, L1, losses[0], losses[-1]),
          flush=True)
    assert L1 < L0

    test = _training_batch([170, 140], [201, 202], kp)
    _, feat_t = _prepare(lp, test)
    rhs = test.right_hand_sides(_plane_waves(4))
    tol = 1e-6
    x_plain, st_plain, ok = test.solve_with_stats(rhs, tol=tol, precision='mixed')
    assert ok
    lp.apply_prior(test.operator)
    x_prior, st_prior, fb_prior = solve_with_fallback(test.operator, test.t_diagonal, rhs, tol)
    lp.apply(test.operator, feat_t)
    x_learn, st_learn, fb_learn = solve_with_fallback(test.operator, test.t_diagonal, rhs, tol)
    print('iterations: plain %d, first-order near-field %d, learned %d; fallbacks %d %d; seconds %.3f %.3f %.3f'
          % (st_plain['iterations'], st_prior['iterations'], st_learn['iterations'], fb_prior, fb_learn,
             st_plain['seconds'], st_prior['seconds'], st_learn['seconds']), flush=True)
    for x, s in ((x_prior, st_prior), (x_learn, st_learn)):
        assert s['relative_residual'] <= tol
        assert np.linalg.norm(x - x_plain) / np.linalg.norm(x_plain) < 20 * tol
    assert not fb_learn
    assert st_learn['iterations'] < st_plain['iterations']
    # validation on a third held-out ensemble picks the network or the first-order structure; the choice
    # is applied to the test ensemble and still beats plain GMRES under the same certificate
    val = _training_batch([160], [301], kp)
    _, feat_v = _prepare(lp, val)
    sel = lp.select(val.operator, feat_v, val.t_diagonal, val.right_hand_sides(_plane_waves(2)), tol)
    print('validation:', sel, flush=True)
    assert sel['choice'] == ('network' if sel['network'] < sel['prior'] else 'prior')
    lp.apply_selected(test.operator, feat_t)
    x_sel, st_sel, fb_sel = solve_with_fallback(test.operator, test.t_diagonal, rhs, tol)
    assert st_sel['relative_residual'] <= tol and not fb_sel
    assert st_sel['iterations'] == (st_learn if sel['choice'] == 'network' else st_prior)['iterations']
    assert st_sel['iterations'] < st_plain['iterations']
