"""Synthetic particle ensembles of the benchmark configurations (seeded).

Random sequential addition with a grid hash: candidate centres are drawn
uniformly and accepted if they keep the minimum separation r_i + r_j + gap
to all accepted spheres.
"""
import numpy as np

from .particles import Sphere


def sequential_addition(n, side_x, side_y, radii, gap, rng, max_tries=200):
    """In-plane centres of n discs with |c_i - c_j| >= r_i + r_j + gap."""
    rmax = float(np.max(radii))
    cell = 2 * rmax + gap
    nx, ny = max(1, int(side_x / cell)), max(1, int(side_y / cell))
    grid = {}
    out = np.zeros((n, 2))
    for i in range(n):
        for _ in range(max_tries * n if i == 0 else max_tries):
            c = np.array([rng.uniform(-side_x / 2, side_x / 2), rng.uniform(-side_y / 2, side_y / 2)])
            gx, gy = int((c[0] + side_x / 2) / cell), int((c[1] + side_y / 2) / cell)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for j in grid.get((gx + dx, gy + dy), ()):
                        if np.hypot(*(c - out[j])) < radii[i] + radii[j] + gap:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                out[i] = c
                grid.setdefault((gx, gy), []).append(i)
                break
        else:
            raise RuntimeError('could not place particle %d of %d' % (i, n))
    del nx, ny
    return out


def monolayer(n, side, z, seed, radius_range=(80.0, 120.0), n_real=(1.5, 2.0), n_imag=0.01, l_max=3, m_max=None,
              gap=20.0):
    """Spheres at one height z, centres in a side x side square (nm)."""
    rng = np.random.default_rng(seed)
    radii = rng.uniform(*radius_range, n)
    nre = rng.uniform(*n_real, n)
    xy = sequential_addition(n, side, side, radii, gap, rng)
    return [Sphere(position=[xy[i, 0], xy[i, 1], z], refractive_index=complex(nre[i], n_imag), radius=float(radii[i]),
                   l_max=l_max, m_max=l_max if m_max is None else m_max) for i in range(n)]


def small_substrate_ensemble(seed=1):
    """50 dielectric spheres (r = 100 nm, n = 1.5 + 0.01i) at z = 200 nm in a 4 x 4 um square,
    220 nm minimum centre separation (the CPU-runnable configuration)."""
    rng = np.random.default_rng(seed)
    radii = np.full(50, 100.0)
    xy = sequential_addition(50, 4000.0, 4000.0, radii, 20.0, rng)
    return [Sphere(position=[x, y, 200.0], refractive_index=1.5 + 0.01j, radius=100.0, l_max=3, m_max=3)
            for x, y in xy]
