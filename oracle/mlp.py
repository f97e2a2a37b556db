"""CPU restatement of the learned near-field preconditioner's network and
training loss (test infrastructure: only tests/, the smoke entry point and the
CPU baseline of bench.py use it).

The reference has no learned component; its GMRES is unpreconditioned
(linearsystem/linear_system.py:271-286). What is restated here is this
project's own definition, so the oracle is pinned by mathematics: its
backward pass against central finite differences of its forward pass (with
the BF16 rounding switched off, in FP64), its loss gradient the same way, and
the BF16 rounding against exactly representable and tie values.

Network: x -> relu(x W1 + b1) -> relu(h1 W2 + b2) -> h2 W3 + b3, parameters
packed as W1 [n_in][n_h], b1, W2 [n_h][n_h], b2, W3 [n_h][n_out], b3. With
round_bf16 the operands of every product (weights, layer inputs,
back-propagated gradients) are rounded to BF16 (round to nearest even) and
the products are accumulated in FP64, as the device accumulates in FP32.

Loss of the preconditioner scales s = 1 + y (y [rows][k1][n_ch][2]):
  E[i, r, n] = E0[i, r, n] + sum_k s[i, k, ch(n)] U[i, k, r, n]
  L = sum |E|^2 / xnorm2 + (w / rows) sum y^2,
  dL/dy[i, k, c] = 2 (Re, Im) sum_{r, ch(n) = c} conj(U) E / xnorm2 + 2 (w / rows) y.
"""
import numpy as np


def bf16(a):
    """Round float32 values to BF16 (round to nearest even), returned as float64."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, n_in, n_h, n_out):
    p = np.asarray(params, dtype=np.float64)
    o = 0
    W1 = p[o:o + n_in * n_h].reshape(n_in, n_h)
    o += n_in * n_h
    b1 = p[o:o + n_h]
    o += n_h
    W2 = p[o:o + n_h * n_h].reshape(n_h, n_h)
    o += n_h * n_h
    b2 = p[o:o + n_h]
    o += n_h
    W3 = p[o:o + n_h * n_out].reshape(n_h, n_out)
    o += n_h * n_out
    b3 = p[o:o + n_out]
    return W1, b1, W2, b2, W3, b3


def _r(a, on):
    return bf16(a) if on else np.asarray(a, dtype=np.float64)


def forward(params, x, n_in, n_h, n_out, round_bf16=True):
    """(y, h1, h2) for rows x [rows][n_in]."""
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, n_h, n_out)
    h1 = np.maximum(_r(x, round_bf16) @ _r(W1, round_bf16) + b1, 0.0)
    h2 = np.maximum(_r(h1, round_bf16) @ _r(W2, round_bf16) + b2, 0.0)
    y = _r(h2, round_bf16) @ _r(W3, round_bf16) + b3
    return y, h1, h2


def backward(params, x, dy, n_in, n_h, n_out, round_bf16=True):
    """Gradient of sum(dy * y) with respect to the packed parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(params, n_in, n_h, n_out)
    y, h1, h2 = forward(params, x, n_in, n_h, n_out, round_bf16)
    r = lambda a: _r(a, round_bf16)  # noqa: E731
    dW3 = r(h2).T @ r(dy)
    db3 = np.asarray(dy, dtype=np.float64).sum(0)
    dh2 = (r(dy) @ r(W3).T) * (h2 > 0)
    dW2 = r(h1).T @ r(dh2)
    db2 = dh2.sum(0)
    dh1 = (r(dh2) @ r(W2).T) * (h1 > 0)
    dW1 = r(x).T @ r(dh1)
    db1 = dh1.sum(0)
    return np.concatenate([dW1.ravel(), db1, dW2.ravel(), db2, dW3.ravel(), db3])


def channels(l_max, m_max):
    """(tau, l) channel tau l_max + l - 1 of every multipole in smuthi's (tau, l, m) order."""
    out = []
    for tau in range(2):
        for l in range(1, l_max + 1):
            for m in range(-min(l, m_max), min(l, m_max) + 1):
                out.append(tau * l_max + l - 1)
    return np.array(out)


def loss(y, U, E0, ch, xnorm2, prior_weight=0.0):
    """(L, dL/dy) for y [rows][k1 n_ch 2], U [rows][k1][nrhs][B], E0 [rows][nrhs][B]; the pull towards
    y = 0 adds (prior_weight / rows) sum y^2."""
    rows, k1, nrhs, B = U.shape
    n_ch = int(ch.max()) + 1
    yc = np.asarray(y, dtype=np.float64).reshape(rows, k1, -1, 2)
    s = 1.0 + yc[..., 0] + 1j * yc[..., 1]  # [rows][k1][n_ch']
    sn = s[:, :, ch]  # [rows][k1][B]
    E = E0 + np.einsum('ikn,ikrn->irn', sn, U)
    L = float(np.sum(np.abs(E) ** 2) / xnorm2)
    g = np.einsum('ikrn,irn->ikn', np.conj(U), E)  # [rows][k1][B]
    gc = np.zeros((rows, k1, yc.shape[2]), dtype=complex)
    for c in range(n_ch):
        gc[:, :, c] = g[:, :, ch == c].sum(-1)
    dy = np.stack([2 * gc.real, 2 * gc.imag], -1) / xnorm2
    yf = np.asarray(y, dtype=np.float64).reshape(rows, -1)
    L += float(prior_weight / rows * np.sum(yf ** 2))
    return L, dy.reshape(rows, -1) + 2 * prior_weight / rows * yf


def near_products(A, nbr, R):
    """U[i, k, r, n] = sum_n2 A[i, k, n, n2] R[r, nbr[i, k], n2] (zero for empty slots);
    A [rows][k1][B][B], R [nrhs][rows][B] (padded)."""
    rows, k1, B, _ = A.shape
    U = np.zeros((rows, k1, R.shape[0], B), dtype=complex)
    for i in range(rows):
        for k in range(k1):
            j = nbr[i, k]
            if j >= 0:
                U[i, k] = (A[i, k] @ R[:, j, :].T).T
    return U


def precondition(A, nbr, s, ch, v):
    """(P v)_i = v_i + sum_k diag(s[i, k, ch]) A[i, k] v_nbr[i, k] for padded v [rows][B]."""
    out = np.array(v, dtype=complex)
    for i in range(A.shape[0]):
        for k in range(A.shape[1]):
            j = nbr[i, k]
            if j >= 0:
                out[i] += s[i, k, ch] * (A[i, k] @ v[j])
    return out
