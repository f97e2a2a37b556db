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


def film_ensemble(n, seed, z_range, radius_range, n_real, n_imag=(0.01, 0.01), l_max=3, density=2.78e-6, gap=20.0,
                  max_tries=400):
    """Spheres at heights inside a film: radii r ~ U(radius_range), Re n ~ U(n_real), Im n ~ U(n_imag), centre
    heights z ~ U(z_range(r)) (z_range: radius -> (low, high)), in-plane density `density` per nm^2 in a
    square centred on the origin, centres at least r_i + r_j + gap apart in 3-D (random sequential
    addition with a grid hash)."""
    rng = np.random.default_rng(seed)
    side = float(np.sqrt(n / density))
    radii = rng.uniform(*radius_range, n)
    nre = rng.uniform(*n_real, n)
    nim = rng.uniform(*n_imag, n)
    cell = 2 * radii.max() + gap
    grid = {}
    pos = np.zeros((n, 3))
    for i in range(n):
        lo, hi = z_range(radii[i])
        for _ in range(max_tries * (n if i == 0 else 1)):
            c = np.array([rng.uniform(-side / 2, side / 2), rng.uniform(-side / 2, side / 2), rng.uniform(lo, hi)])
            gx, gy = int((c[0] + side / 2) // cell), int((c[1] + side / 2) // cell)
            ok = True
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for j in grid.get((gx + dx, gy + dy), ()):
                        if np.linalg.norm(c - pos[j]) < radii[i] + radii[j] + gap:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                pos[i] = c
                grid.setdefault((gx, gy), []).append(i)
                break
        else:
            raise RuntimeError('could not place particle %d of %d' % (i, n))
    return [Sphere(position=list(pos[i]), refractive_index=complex(nre[i], nim[i]), radius=float(radii[i]),
                   l_max=l_max, m_max=l_max) for i in range(n)]


def batching_ensemble(n, seed):
    """Batching configuration: r ~ U(80, 120) nm, n ~ U(1.5, 2.0) + 0.01i, heights ~ U(150, 600) nm inside
    a 1 um film (n = 1.8) on glass, 2.78 spheres per um^2, l_max = 3."""
    return film_ensemble(n, seed, lambda r: (150.0, 600.0), (80.0, 120.0), (1.5, 2.0))


def batching_sizes(count=32, seed=4, low=1000, high=20000):
    """Ensemble sizes round(exp(U(ln low, ln high)))."""
    rng = np.random.default_rng(seed)
    return [int(round(np.exp(v))) for v in rng.uniform(np.log(low), np.log(high), count)]


def learned_ensemble(n, seed):
    """Learned-preconditioner configuration: r ~ U(60, 180) nm, Re n ~ U(1.4, 2.4), Im n ~ U(0, 0.05),
    centres inside the 500 nm film (n = 2.0) of air / film / glass at z ~ U(r + 10, 490 - r) nm,
    2.78 spheres per um^2, l_max = 3."""
    return film_ensemble(n, seed, lambda r: (r + 10.0, 490.0 - r), (60.0, 180.0), (1.4, 2.4), (0.0, 0.05))
