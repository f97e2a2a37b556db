"""GPU parity of the direct spherical-wave-expansion field (smh_swe_field:
SphericalWaveExpansion.electric_field / magnetic_field, fields/expansions.py:348-424,
summed over the particles as PiecewiseFieldExpansion does,
postprocessing/scattered_field.py:72-75) against the FP64 oracle."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import swe as osw
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _particles(n, seed):
    rng = np.random.default_rng(seed)
    orders = [(3, 3), (3, 2), (2, 2), (4, 3), (1, 1)]
    plist = []
    for i in range(n):
        l, m = orders[i % len(orders)]
        p = osys.Sphere([rng.uniform(-1500, 1500), rng.uniform(-1500, 1500), rng.uniform(80, 320)], 1.6, 60, l, m)
        p.scattered_field = (rng.standard_normal(ob.blocksize(l, m)) + 1j * rng.standard_normal(ob.blocksize(l, m)))
        plist.append(p)
    return plist


def _points(plist, n, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1800, 1800, n)
    y = rng.uniform(-1800, 1800, n)
    z = rng.uniform(-100, 500, n)
    # a particle centre (inside), points just outside two spheres, points near the layer bounds
    extra = [plist[0].position, np.add(plist[1].position, [0, 0, 60.5]), np.add(plist[2].position, [61, 0, 0]),
             [0.0, 0.0, 0.0], [10.0, 20.0, 399.9], [10.0, 20.0, 400.0]]
    e = np.array(extra)
    return np.concatenate([x, e[:, 0]]), np.concatenate([y, e[:, 1]]), np.concatenate([z, e[:, 2]])


def _rel(got, want):
    scale = max(np.abs(w).max() for w in want)
    return max(np.abs(g - w).max() for g, w in zip(got, want)) / scale


@pytest.mark.parametrize('field', ['E', 'H'])
def test_swe_sum_matches_oracle(gpu, field):
    from smuthi_amd.expansions import SphericalWaveSum
    plist = _particles(40, 3)
    k = 2 * np.pi / WL * 1.5 + 0.003j
    x, y, z = _points(plist, 700, 4)
    want = osw.direct_field([(p.scattered_field, p.l_max, p.m_max, p.position, p.radius) for p in plist], k, x, y, z,
                            field, ob.angular_frequency(WL), z_range=(0.0, 400.0))
    s = SphericalWaveSum(k, plist, 0.0, 400.0)
    for prec, tol in (('double', 1e-11), ('single', 2e-5)):
        got = s.electric_field(x, y, z, precision=prec) if field == 'E' else \
            s.magnetic_field(x, y, z, WL, precision=prec)
        assert _rel(got, want) < tol, (prec, _rel(got, want))
        assert all(np.abs(g[-1]) == 0 for g in got)  # z = upper bound: outside the layer
        # at the centre of particle 0 only the other particles contribute, finite
        assert all(np.isfinite(g[700]) for g in got)


@pytest.mark.parametrize('field', ['E', 'H'])
def test_swe_sum_high_orders(gpu, field):
    # l_max 5..7 and m_max 0: the runtime-order kernel (the unrolled ones take l_max <= 4, m_max >= 1)
    from smuthi_amd.expansions import SphericalWaveSum
    rng = np.random.default_rng(11)
    plist = []
    for i, (l, m) in enumerate([(7, 7), (6, 4), (5, 5), (2, 0), (3, 3), (7, 2)] * 5):
        p = osys.Sphere([rng.uniform(-1500, 1500), rng.uniform(-1500, 1500), rng.uniform(80, 320)], 1.6, 60, l, m)
        p.scattered_field = (rng.standard_normal(ob.blocksize(l, m)) + 1j * rng.standard_normal(ob.blocksize(l, m)))
        plist.append(p)
    k = 2 * np.pi / WL * 1.5 + 0.003j
    x, y, z = _points(plist, 500, 12)
    want = osw.direct_field([(p.scattered_field, p.l_max, p.m_max, p.position, p.radius) for p in plist], k, x, y, z,
                            field, ob.angular_frequency(WL), z_range=(0.0, 400.0))
    s = SphericalWaveSum(k, plist, 0.0, 400.0)
    for prec, tol in (('double', 1e-11), ('single', 5e-5)):
        got = s.electric_field(x, y, z, precision=prec) if field == 'E' else \
            s.magnetic_field(x, y, z, WL, precision=prec)
        assert _rel(got, want) < tol, (prec, _rel(got, want))
    # m_max = 0 alone (a single order group)
    q = [osys.Sphere([0.0, 0.0, 200.0], 1.6, 60, 3, 0)]
    q[0].scattered_field = rng.standard_normal(ob.blocksize(3, 0)) + 1j * rng.standard_normal(ob.blocksize(3, 0))
    want0 = osw.direct_field([(q[0].scattered_field, 3, 0, q[0].position, q[0].radius)], k, x, y, z, field,
                             ob.angular_frequency(WL), z_range=(0.0, 400.0))
    s0 = SphericalWaveSum(k, q, 0.0, 400.0)
    got0 = s0.electric_field(x, y, z, precision='double') if field == 'E' else \
        s0.magnetic_field(x, y, z, WL, precision='double')
    assert _rel(got0, want0) < 1e-11


def test_single_swe_drop_in(gpu):
    from smuthi_amd.expansions import SphericalWaveExpansion
    rng = np.random.default_rng(6)
    k = 2 * np.pi / WL * 1.33
    pos = [100.0, -50.0, 200.0]
    swe = SphericalWaveExpansion(k, 3, 2, 'outgoing', pos, 0.0, 1000.0)
    swe.coefficients = rng.standard_normal(ob.blocksize(3, 2)) + 1j * rng.standard_normal(ob.blocksize(3, 2))
    sph = osys.Sphere(pos, 1.5, 80, 3, 2)
    swe.validity_conditions.append(sph.is_outside)
    x = rng.uniform(-600, 600, 300)
    y = rng.uniform(-600, 600, 300)
    z = rng.uniform(-200, 1200, 300)
    vld = (z >= 0) & (z < 1000) & sph.is_outside(x, y, z)
    we = osw.swe_electric_field(swe.coefficients, 3, 2, k, pos, x, y, z)
    wh = osw.swe_magnetic_field(swe.coefficients, 3, 2, k, pos, x, y, z, ob.angular_frequency(WL))
    ge = swe.electric_field(x, y, z, precision='double')
    gh = swe.magnetic_field(x, y, z, WL, precision='double')
    for got, want in ((ge, we), (gh, wh)):
        want = [np.where(vld, w, 0) for w in want]
        assert _rel(got, want) < 1e-11
