"""The CPU restatement of the learned preconditioner's network and loss,
pinned by mathematics: backward passes against central finite differences of
the forward passes (BF16 rounding off, FP64), the loss gradient likewise, the
BF16 rounding on exact and tie values, and the preconditioner's first-order
structure against the exact inverse of a weakly coupled system."""
import numpy as np

from oracle import mlp


def test_bf16_rounding():
    x = np.array([1.0, -2.5, 3.0e-3, 65504.0], dtype=np.float32)
    # values with at most 8 significant bits are exact
    assert np.array_equal(mlp.bf16(np.array([1.0, -2.5, 0.15625, 384.0], np.float32)), [1.0, -2.5, 0.15625, 384.0])
    # ties go to the even mantissa: 1 + 2^-8 -> 1, 1 + 3 * 2^-8 -> 1 + 2^-6
    assert mlp.bf16(np.array([1 + 2 ** -8], np.float32))[0] == 1.0
    assert mlp.bf16(np.array([1 + 3 * 2 ** -8], np.float32))[0] == 1 + 2 ** -6
    # relative error at most the unit roundoff 2^-8 of an 8-bit significand
    y = mlp.bf16(x)
    assert np.all(np.abs(y - x) <= 2 ** -8 * np.abs(x))


def _params(n_in, n_h, n_out, rng):
    n = n_in * n_h + n_h + n_h * n_h + n_h + n_h * n_out + n_out
    return rng.standard_normal(n) * 0.3


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    n_in, n_h, n_out = 7, 9, 5
    p = _params(n_in, n_h, n_out, rng)
    x = rng.standard_normal((6, n_in))
    dy = rng.standard_normal((6, n_out))
    g = mlp.backward(p, x, dy, n_in, n_h, n_out, round_bf16=False)
    f = lambda q: float(np.sum(dy * mlp.forward(q, x, n_in, n_h, n_out, round_bf16=False)[0]))  # noqa: E731
    eps = 1e-6
    idx = rng.choice(len(p), 60, replace=False)
    for k in idx:
        e = np.zeros_like(p)
        e[k] = eps
        fd = (f(p + e) - f(p - e)) / (2 * eps)
        assert abs(fd - g[k]) <= 1e-6 * (1 + abs(g[k])), (k, fd, g[k])


def test_loss_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    rows, k1, nrhs, l_max = 4, 3, 2, 2
    ch = mlp.channels(l_max, l_max)
    B = len(ch)
    n_ch = 2 * l_max
    U = rng.standard_normal((rows, k1, nrhs, B)) + 1j * rng.standard_normal((rows, k1, nrhs, B))
    E0 = rng.standard_normal((rows, nrhs, B)) + 1j * rng.standard_normal((rows, nrhs, B))
    y = 0.1 * rng.standard_normal((rows, k1 * n_ch * 2))
    for w in (0.0, 0.3):
        L, dy = mlp.loss(y, U, E0, ch, 7.0, w)
        eps = 1e-6
        for i in range(rows):
            for k in range(0, dy.shape[1], 3):
                e = np.zeros_like(y)
                e[i, k] = eps
                fd = (mlp.loss(y + e, U, E0, ch, 7.0, w)[0] - mlp.loss(y - e, U, E0, ch, 7.0, w)[0]) / (2 * eps)
                assert abs(fd - dy[i, k]) <= 1e-6 * (1 + abs(dy[i, k])), (w, i, k, fd, dy[i, k])
    # the pull's own value: y = 0 except one entry
    y0 = np.zeros_like(y)
    y0[1, 4] = 0.5
    L0 = mlp.loss(np.zeros_like(y), U, E0, ch, 7.0, 0.3)[0]
    L1 = mlp.loss(y0, U, E0, ch, 7.0, 0.3)[0] - mlp.loss(y0, U, E0, ch, 7.0, 0.0)[0]
    assert abs(L1 - 0.3 / rows * 0.25) < 1e-12 and L0 == mlp.loss(np.zeros_like(y), U, E0, ch, 7.0)[0]


def test_first_order_preconditioner_inverts_weak_coupling():
    """s = 1 makes P = 1 + A_near: for a system whose coupling is only between near pairs and weak
    (|A| ~ eps), M P = (1 - A)(1 + A) = 1 - A^2, i.e. P is the inverse to second order."""
    rng = np.random.default_rng(2)
    rows, B, k1 = 5, 6, 3
    ch = np.arange(B) % 2
    nbr = np.array([[i, (i + 1) % rows, (i + 2) % rows] for i in range(rows)])
    eps = 1e-3
    A = eps * (rng.standard_normal((rows, k1, B, B)) + 1j * rng.standard_normal((rows, k1, B, B)))
    Md = np.eye(rows * B, dtype=complex)
    for i in range(rows):
        for k in range(k1):
            j = nbr[i, k]
            Md[i * B:(i + 1) * B, j * B:(j + 1) * B] -= A[i, k]
    s = np.ones((rows, k1, 2), dtype=complex)
    v = rng.standard_normal((rows, B)) + 1j * rng.standard_normal((rows, B))
    pv = mlp.precondition(A, nbr, s, ch, v).ravel()
    res = Md @ pv - v.ravel()
    assert np.linalg.norm(res) / np.linalg.norm(v) < 50 * eps ** 2
