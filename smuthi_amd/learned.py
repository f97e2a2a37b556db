"""Learned near-field preconditioner for the batched coupling solve.

No counterpart in smuthi (its GMRES is unpreconditioned,
linear_system.py:271-286). The right preconditioner of particle i corrects
the T-matrix-scaled system M = 1 - T W over its K nearest neighbours N(i)
(slot 0 is i itself):

    (P v)_i = v_i + sum_{j in {i} + N(i)} diag(s_ij[ch]) A_ij v_j,   A_ij = diag(t_i) W_ij,

A_ij being the coupling block the operator's lookup applies (built on the
device, smh_op_near_blocks) and s_ij one complex scale per (tau, l) channel
of the receiving multipole. With s = 1 this is the first-order near-field
inverse of M; a 3-layer network (hidden width 256, BF16 weights, FP32
accumulation, hand-written HIP forward / backward / Adam, smh_mlp_*) maps
each particle's neighbourhood -- positions of its K = 16 nearest neighbours
relative to it, their radii and refractive indices, and its place in the
layer -- to s - 1. It is trained on solutions the solver generates (random
and plane-wave right-hand sides, FP64-certified) by fitting P to the
solution operator, L = sum |P R - X|^2 / sum |X|^2, with an optional pull
of the scales towards s = 1. A validation solve (select) keeps the network
only where it beats the first-order structure.

Plain GMRES stays both judge and fallback: every solve is certified by the
FP64 residual ||b - M x|| <= tol ||b|| of the reference's stopping rule, and
a preconditioned solve that misses it within its budget is re-run without
the preconditioner (solve_with_fallback).
"""
import ctypes

import numpy as np
import scipy.spatial

from . import _lib
from . import cuda as cu

N_SELF = 8  # features of the particle itself
N_PER_NEIGHBOUR = 8


def feature_count(K):
    return N_SELF + N_PER_NEIGHBOUR * K


def neighbourhoods(positions, radii, refractive_indices, layer_bottom, layer_top, wavelength, K=16,
                   ensemble_offsets=None):
    """(nbr [n, K + 1] int32, features [n, 8 + 8 K] float32) of every particle.

    nbr[i, 0] = i, then the K nearest particles of i's ensemble (3-D centre
    distance, nearest first; -1 where an ensemble has fewer). Features, in
    units of the wavelength lambda, rotation-invariant about the layer normal:
      self: radius / 100 nm, Re n, 10 Im n, height above the layer's bottom /
            lambda, depth below its top / lambda (both clipped to 4), distance
            to the K-th neighbour / lambda, neighbours within lambda / K, 1;
      per neighbour: in-plane distance / lambda, z difference / lambda,
            distance / lambda, lambda / (2 pi distance), (z_i + z_j - 2 bottom)
            / lambda (clipped to 8), radius / 100 nm, Re n, 10 Im n (zeros for
            an empty slot)."""
    pos = np.asarray(positions, dtype=float)
    n = len(pos)
    r = np.asarray(radii, dtype=float)
    nr = np.asarray(refractive_indices, dtype=complex)
    off = np.array([0, n] if ensemble_offsets is None else ensemble_offsets, dtype=np.int64)
    lam = float(wavelength)
    nbr = -np.ones((n, K + 1), dtype=np.int32)
    nbr[:, 0] = np.arange(n)
    dist = np.zeros((n, K))
    for a, b in zip(off[:-1], off[1:]):
        m = int(b - a)
        if m < 2:
            continue
        kk = min(K, m - 1)
        d, idx = scipy.spatial.cKDTree(pos[a:b]).query(pos[a:b], k=kk + 1)
        nbr[a:b, 1:kk + 1] = (idx[:, 1:] + a).astype(np.int32)
        dist[a:b, :kk] = d[:, 1:]
    bottom = layer_bottom if np.isfinite(layer_bottom) else pos[:, 2].min() - 4 * lam
    top = layer_top if np.isfinite(layer_top) else pos[:, 2].max() + 4 * lam
    feat = np.zeros((n, feature_count(K)), dtype=np.float32)
    z = pos[:, 2]
    valid = nbr[:, 1:] >= 0
    feat[:, 0] = r / 100.0
    feat[:, 1] = nr.real
    feat[:, 2] = 10 * nr.imag
    feat[:, 3] = np.minimum((z - bottom) / lam, 4.0)
    feat[:, 4] = np.minimum((top - z) / lam, 4.0)
    feat[:, 5] = np.where(valid[:, -1], dist[:, -1], 0.0) / lam
    feat[:, 6] = np.sum((dist < lam) & valid, axis=1) / K
    feat[:, 7] = 1.0
    j = np.where(valid, nbr[:, 1:], 0)
    dx = pos[j, 0] - pos[:, None, 0]
    dy = pos[j, 1] - pos[:, None, 1]
    dz = pos[j, 2] - pos[:, None, 2]
    rho = np.hypot(dx, dy)
    d = np.sqrt(rho ** 2 + dz ** 2)
    blk = np.stack([rho / lam, dz / lam, d / lam, lam / (2 * np.pi * np.maximum(d, 1e-9)),
                    np.minimum((pos[:, None, 2] + pos[j, 2] - 2 * bottom) / lam, 8.0), r[j] / 100.0, nr[j].real,
                    10 * nr[j].imag], axis=-1)
    blk[~valid] = 0.0
    feat[:, N_SELF:] = blk.reshape(n, -1).astype(np.float32)
    return nbr, feat


class LearnedPreconditioner:
    """The network and its training set on the device (smh_mlp). K nearest
    neighbours, hidden width `hidden`, outputs (K + 1) x 2 l_max complex
    scales. The output layer starts at zero: the untrained network gives the
    first-order near-field inverse (s = 1)."""

    def __init__(self, K=16, hidden=256, l_max=3, seed=0, prior_weight=0.0):
        self.K, self.hidden, self.l_max = K, hidden, l_max
        self.n_in = feature_count(K)
        self.n_ch = 2 * l_max
        self.n_out = (K + 1) * self.n_ch * 2
        handle = ctypes.c_void_p()
        _lib.check(_lib.load().smh_mlp_create(cu.context(), self.n_in, hidden, self.n_out, seed, ctypes.byref(handle)),
                   cu.context(), 'smh_mlp_create')
        self._mlp = handle
        self.losses = []
        self.choice = 'network'
        if prior_weight:
            self.set_prior_weight(prior_weight)

    def set_prior_weight(self, w):
        """Weight w of the pull of the scales towards the first-order structure: the training loss adds
        (w / rows) sum |y|^2 (s = 1 + y)."""
        _lib.check(_lib.load().smh_mlp_set_prior_weight(self._mlp, float(w)), cu.context(),
                   'smh_mlp_set_prior_weight')

    def __del__(self):
        if getattr(self, '_mlp', None) is not None and _lib._LIB is not None:
            _lib.load().smh_mlp_destroy(self._mlp)
            self._mlp = None

    @property
    def n_params(self):
        return int(_lib.load().smh_mlp_num_params(self._mlp))

    def params(self):
        out = np.empty(self.n_params, dtype=np.float32)
        _lib.check(_lib.load().smh_mlp_get_params(self._mlp, _lib.fptr(out)), cu.context(), 'smh_mlp_get_params')
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.shape != (self.n_params,):
            raise ValueError('expected %d parameters' % self.n_params)
        _lib.check(_lib.load().smh_mlp_set_params(self._mlp, _lib.fptr(p)), cu.context(), 'smh_mlp_set_params')

    def forward(self, feat):
        x = np.ascontiguousarray(feat, dtype=np.float32)
        y = np.empty((len(x), self.n_out), dtype=np.float32)
        _lib.check(_lib.load().smh_mlp_forward(self._mlp, len(x), _lib.fptr(x), _lib.fptr(y)), cu.context(),
                   'smh_mlp_forward')
        return y

    def backward(self, feat, dy):
        x = np.ascontiguousarray(feat, dtype=np.float32)
        d = np.ascontiguousarray(dy, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        _lib.check(_lib.load().smh_mlp_backward(self._mlp, len(x), _lib.fptr(x), _lib.fptr(d), _lib.fptr(g)),
                   cu.context(), 'smh_mlp_backward')
        return g

    def prepare(self, operator, positions, radii, refractive_indices, layer_bottom, layer_top, wavelength,
                t_diagonal, ensemble_offsets=None):
        """Neighbourhoods of the operator's particles and their near blocks on the device: returns
        (nbr, features)."""
        nbr, feat = neighbourhoods(positions, radii, refractive_indices, layer_bottom, layer_top, wavelength, self.K,
                                   ensemble_offsets)
        nb = np.ascontiguousarray(nbr)
        t = np.ascontiguousarray(t_diagonal, dtype=np.complex128)
        _lib.check(_lib.load().smh_op_near_blocks(operator._op, self.K + 1, _lib.iptr(nb), _lib.dptr(t)),
                   cu.context(), 'smh_op_near_blocks')
        return nbr, feat

    def add_training_ensemble(self, operator, feat, rhs, sol):
        """Add an ensemble's particles (near blocks prepared) with right-hand sides and the solver's
        solutions [nrhs, n_unk] to the training set."""
        x = np.ascontiguousarray(feat, dtype=np.float32)
        r = np.ascontiguousarray(rhs, dtype=np.complex128)
        s = np.ascontiguousarray(sol, dtype=np.complex128)
        _lib.check(_lib.load().smh_mlp_add_data(self._mlp, operator._op, _lib.fptr(x), len(r), _lib.dptr(r),
                                                _lib.dptr(s)), cu.context(), 'smh_mlp_add_data')

    def data_loss(self, y=None):
        """(L, dL/dy) over the training set at outputs y (default: the network's)."""
        loss = ctypes.c_double()
        rows = None
        yy = None if y is None else np.ascontiguousarray(y, dtype=np.float32)
        if yy is not None:
            rows = len(yy)
        dy = None if rows is None else np.empty_like(yy)
        _lib.check(_lib.load().smh_mlp_data_loss(self._mlp, _lib.fptr(yy), _lib.fptr(dy), ctypes.byref(loss)),
                   cu.context(), 'smh_mlp_data_loss')
        return loss.value, dy

    def train(self, steps, lr=1e-3):
        """Full-batch Adam steps on the device; returns the loss before each step."""
        out = np.zeros(steps)
        _lib.check(_lib.load().smh_mlp_train(self._mlp, int(steps), float(lr), _lib.dptr(out)), cu.context(),
                   'smh_mlp_train')
        self.losses.extend(out.tolist())
        return out

    def scales(self, feat):
        """s = 1 + network(features) as float32 [n, K + 1, 2 l_max, 2]."""
        y = self.forward(feat).reshape(len(feat), self.K + 1, self.n_ch, 2)
        y[..., 0] += 1.0
        return np.ascontiguousarray(y)

    def apply(self, operator, feat):
        """Set the learned preconditioner on the operator (near blocks prepared)."""
        s = self.scales(feat)
        _lib.check(_lib.load().smh_op_set_near_scaling(operator._op, _lib.fptr(s)), cu.context(),
                   'smh_op_set_near_scaling')

    def apply_prior(self, operator):
        """The untrained structure, s = 1 (first-order near-field inverse)."""
        n = len(operator.particle_list)
        s = np.zeros((n, self.K + 1, self.n_ch, 2), dtype=np.float32)
        s[..., 0] = 1.0
        _lib.check(_lib.load().smh_op_set_near_scaling(operator._op, _lib.fptr(s)), cu.context(),
                   'smh_op_set_near_scaling')

    def select(self, operator, feat, t_diagonal, rhs, tol, precision='mixed'):
        """Validation on held-out systems: solve rhs with the network's scales and with the first-order
        structure (s = 1) and keep the one that needs fewer GMRES iterations (the structure on a tie or when
        the network's solve falls back). Returns the iterations of both and the choice."""
        self.apply(operator, feat)
        _, st_n, fb_n = solve_with_fallback(operator, t_diagonal, rhs, tol, precision=precision)
        self.apply_prior(operator)
        _, st_p, fb_p = solve_with_fallback(operator, t_diagonal, rhs, tol, precision=precision)
        self.remove(operator)
        net = st_n['iterations'] if not fb_n else np.inf
        pri = st_p['iterations'] if not fb_p else np.inf
        self.choice = 'network' if net < pri else 'prior'
        return {'network': st_n['iterations'], 'prior': st_p['iterations'], 'network_fell_back': fb_n,
                'prior_fell_back': fb_p, 'choice': self.choice}

    def apply_selected(self, operator, feat):
        """Set the preconditioner that won the last validation (select): the network's or the first-order
        structure."""
        if self.choice == 'network':
            self.apply(operator, feat)
        else:
            self.apply_prior(operator)

    @staticmethod
    def remove(operator):
        _lib.check(_lib.load().smh_op_set_near_scaling(operator._op, None), cu.context(), 'smh_op_set_near_scaling')


def solve_with_fallback(operator, t_diagonal, rhs, tol, precision='mixed', restart=20, maxiter=None):
    """Solve with the operator's preconditioner; if that misses the FP64 tolerance within the budget, remove
    it and solve with plain GMRES (the fallback). Returns (x, stats, fell_back); stats of the run that
    produced x. The preconditioner is left removed after a fallback."""
    x, st, ok = operator.solve_master(t_diagonal, rhs, tol, precision=precision, restart=restart, maxiter=maxiter,
                                      with_stats=True)
    if ok:
        return x, st, False
    LearnedPreconditioner.remove(operator)
    x, st, ok = operator.solve_master(t_diagonal, rhs, tol, precision=precision, restart=restart, maxiter=maxiter,
                                      with_stats=True)
    if not ok:
        raise RuntimeError('plain GMRES missed the tolerance too: %s' % st)
    return x, st, True
