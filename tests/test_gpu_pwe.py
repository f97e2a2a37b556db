"""GPU parity of the plane-wave-expansion field kernels (PlaneWaveExpansion.
electric_field / magnetic_field, fields/expansions.py:635-782) against the FP64
oracle's nested trapezoid sums."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _pair(kind, k, kp, alpha, g, ref, lower_z=-np.inf, upper_z=np.inf):
    from smuthi_amd.expansions import PlaneWaveExpansion
    o = osys.PlaneWaveExpansion(k, kp, alpha, kind, ref, lower_z, upper_z)
    o.coefficients = g.copy()
    p = PlaneWaveExpansion(k, kp, alpha, kind, ref, lower_z, upper_z)
    p.coefficients = g.copy()
    return o, p


def _rel(got, want):
    scale = max(np.abs(w).max() for w in want)
    return max(np.abs(a - b).max() for a, b in zip(got, want)) / scale


def test_discrete_plane_wave_closed_form(gpu):
    k = 2 * np.pi / WL * 1.5
    th, al = 0.4, 0.3
    g = np.zeros((2, 1, 1), complex)
    g[0, 0, 0], g[1, 0, 0] = 1.0, 0.5j
    o, p = _pair('upgoing', k, [k * np.sin(th)], [al], g, [10.0, -20.0, 30.0])
    rng = np.random.default_rng(0)
    x, y, z = rng.uniform(-3000, 3000, (3, 300))
    want = o.electric_field(x, y, z)
    # closed form: g_TE (-sin a, cos a, 0) + g_TM (cos a kz/k, sin a kz/k, -kp/k), times e^{i k.r}
    kz = k * np.cos(th)
    ph = np.exp(1j * (k * np.sin(th) * (np.cos(al) * (x - 10) + np.sin(al) * (y + 20)) + kz * (z - 30)))
    closed = [(-np.sin(al) * g[0, 0, 0] + np.cos(al) * kz / k * g[1, 0, 0]) * ph,
              (np.cos(al) * g[0, 0, 0] + np.sin(al) * kz / k * g[1, 0, 0]) * ph,
              -np.sin(th) * g[1, 0, 0] * ph]
    assert _rel(want, closed) < 1e-12
    assert _rel(p.electric_field(x, y, z, precision='double'), closed) < 1e-12
    # FP32: phase error ~ 6e-8 * |k| * (leaf extent); these sparse points make leaves of several micrometres
    assert _rel(p.electric_field(x, y, z), closed) < 5e-5


@pytest.mark.parametrize('kind', ['upgoing', 'downgoing'])
def test_continuous_pwe_matches_oracle(gpu, kind):
    k = 2 * np.pi / WL
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.0, 1.0], neff_resolution=2e-2,
                                               neff_imag=1e-2, neff_max=1.6)
    alpha = np.linspace(0, 2 * np.pi, 24)
    rng = np.random.default_rng(3 if kind == 'upgoing' else 4)
    g = (rng.standard_normal((2, len(kp), len(alpha))) + 1j * rng.standard_normal((2, len(kp), len(alpha)))) * 1e-3
    ref = [100.0, -50.0, 0.0]
    lower, upper = (0.0, np.inf) if kind == 'upgoing' else (-np.inf, 0.0)
    o, p = _pair(kind, k, kp, alpha, g, ref, lower, upper)
    n = 700  # several leaves of 256 points
    x, y = rng.uniform(-2500, 2500, (2, n))
    z = rng.uniform(20, 600, n) * (1 if kind == 'upgoing' else -1)
    z[:5] = -z[:5]  # a few points outside the expansion's validity
    wantE = o.electric_field(x, y, z)
    wantH = o.magnetic_field(x, y, z, WL)
    gotE64 = p.electric_field(x, y, z, precision='double')
    gotH64 = p.magnetic_field(x, y, z, WL, precision='double')
    assert _rel(gotE64, wantE) < 1e-10
    assert _rel(gotH64, wantH) < 1e-10
    gotE = p.electric_field(x, y, z)
    gotH = p.magnetic_field(x, y, z, WL)
    assert _rel(gotE, wantE) < 5e-5
    assert _rel(gotH, wantH) < 5e-5
    for c in range(3):  # zero outside the validity range
        assert np.all(gotE[c][:5] == 0) and np.all(wantE[c][:5] == 0)


def test_far_points_keep_fp32_accuracy(gpu):
    # points tens of micrometres from the reference point: the leaf origin
    # carries the large phase in FP64, the per-term phase stays small
    k = 2 * np.pi / WL * 1.52
    kp = np.linspace(0.0, 1.4 * k, 40) + 0j
    alpha = np.linspace(0, 2 * np.pi, 16)
    rng = np.random.default_rng(5)
    g = rng.standard_normal((2, len(kp), len(alpha))) + 1j * rng.standard_normal((2, len(kp), len(alpha)))
    o, p = _pair('downgoing', k, kp, alpha, g, [0.0, 0.0, 0.0], -np.inf, 0.0)
    x = rng.uniform(-30000, 30000, 300)
    y = np.zeros(300)
    z = rng.uniform(-2000, -10, 300)
    want = o.electric_field(x, y, z)
    assert _rel(p.electric_field(x, y, z), want) < 5e-5
