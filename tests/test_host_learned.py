"""Host part of the learned preconditioner: neighbourhoods and features
(no device needed)."""
import numpy as np

from smuthi_amd import learned


def _cloud(n, rng):
    pos = np.column_stack([rng.uniform(-2000, 2000, n), rng.uniform(-2000, 2000, n), rng.uniform(100, 400, n)])
    return pos, rng.uniform(60, 180, n), rng.uniform(1.4, 2.4, n) + 1j * rng.uniform(0, 0.05, n)


def test_neighbours_are_the_nearest_within_each_ensemble():
    rng = np.random.default_rng(0)
    pos, r, n_ = _cloud(90, rng)
    off = [0, 40, 43, 90]  # an ensemble with fewer particles than K
    nbr, feat = learned.neighbourhoods(pos, r, n_, 0.0, 500.0, 550.0, K=16, ensemble_offsets=off)
    assert nbr.shape == (90, 17) and feat.shape == (90, learned.feature_count(16))
    for a, b in zip(off[:-1], off[1:]):
        for i in range(a, b):
            assert nbr[i, 0] == i
            d = np.linalg.norm(pos[a:b] - pos[i], axis=1)
            order = [j + a for j in np.argsort(d) if j + a != i][:16]
            got = [j for j in nbr[i, 1:] if j >= 0]
            assert got == order
            assert all(a <= j < b for j in got)
    # 3-particle ensemble: two neighbours, the rest empty with zero features
    assert (nbr[40:43, 3:] == -1).all()
    assert not feat[40:43, learned.N_SELF + 2 * learned.N_PER_NEIGHBOUR:].any()


def test_features_are_rotation_invariant():
    rng = np.random.default_rng(1)
    pos, r, n_ = _cloud(60, rng)
    c, s = np.cos(0.7), np.sin(0.7)
    rot = pos @ np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])
    nb1, f1 = learned.neighbourhoods(pos, r, n_, 0.0, 500.0, 550.0, K=16)
    nb2, f2 = learned.neighbourhoods(rot, r, n_, 0.0, 500.0, 550.0, K=16)
    assert np.array_equal(nb1, nb2)
    assert np.allclose(f1, f2, atol=1e-5)
