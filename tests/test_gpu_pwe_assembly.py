"""GPU parity of the scattered-field plane wave expansion (the particle sum of
swe_to_pwe_conversion, fields/transformations.py:133-226, in
scattered_field_pwe, postprocessing/scattered_field.py:80-150) and of the
extinction cross section (postprocessing/far_field.py:517-654) against the
FP64 oracle."""
import numpy as np
import pytest

from oracle import basics as ob
from oracle import layers as ol
from oracle import system as osys

pytestmark = pytest.mark.gpu

WL = 550.0


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _particles(n, seed):
    rng = np.random.default_rng(seed)
    plist = []
    orders = [(3, 3), (3, 2), (2, 2), (4, 3)]
    for i in range(n):
        l, m = orders[i % len(orders)]
        p = osys.Sphere([rng.uniform(-1500, 1500), rng.uniform(-1500, 1500), rng.uniform(80, 320)], 1.6, 60, l, m)
        nb = 2 * (m * (m + 2) + (l - m) * (2 * m + 1))
        p.scattered_field = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)) * rng.uniform(0.1, 1)
        plist.append(p)
    return plist


@pytest.mark.parametrize('layer_number,direct,mediated', [(1, True, True), (0, True, True), (2, True, True),
                                                          (1, True, False), (1, False, True)])
def test_scattered_field_pwe_matches_oracle(gpu, layer_number, direct, mediated):
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.scattered_field import scattered_field_pwe
    thick, n = [0, 400, 0], [1.5, 2.0 + 0.01j, 1.0]
    plist = _particles(150, 7)  # > one leaf of 128 particles
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=n, neff_resolution=5e-2,
                                               neff_imag=1e-2)
    alpha = np.linspace(0, 2 * np.pi, 12)
    ou, od = osys.scattered_field_pwe(WL, plist, ol.LayerSystem(thick, n), layer_number, kp, alpha,
                                      include_direct=direct, include_layer_response=mediated)
    for prec, tol in (('double', 1e-10), ('single', 2e-5)):
        pu, pd = scattered_field_pwe(WL, plist, LayerSystem(thick, n), layer_number, kp, alpha,
                                     include_direct=direct, include_layer_response=mediated, precision=prec)
        assert pu.lower_z == ou.lower_z and pu.upper_z == ou.upper_z
        assert pd.lower_z == od.lower_z and pd.upper_z == od.upper_z
        # one scale for both directions: in a half-space one of them vanishes up to rounding
        scale = max(np.abs(ou.coefficients).max(), np.abs(od.coefficients).max())
        assert np.abs(pu.coefficients - ou.coefficients).max() < tol * scale, prec
        assert np.abs(pd.coefficients - od.coefficients).max() < tol * scale, prec


def test_scattered_field_pwe_high_orders(gpu):
    # m_max 5..7: more than 9 orders m, assembled in several launches that add into one expansion;
    # mixed cut-offs (particles padded to the largest l_max, m_max)
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.scattered_field import scattered_field_pwe
    rng = np.random.default_rng(21)
    plist = []
    for i, (l, m) in enumerate([(7, 7), (6, 5), (5, 5), (3, 2), (7, 4), (6, 6)] * 25):
        p = osys.Sphere([rng.uniform(-1500, 1500), rng.uniform(-1500, 1500), rng.uniform(80, 320)], 1.6, 60, l, m)
        nb = 2 * (m * (m + 2) + (l - m) * (2 * m + 1))
        p.scattered_field = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)) * rng.uniform(0.1, 1)
        plist.append(p)
    thick, n = [0, 400, 0], [1.5, 2.0 + 0.01j, 1.0]
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=n, neff_resolution=5e-2,
                                               neff_imag=1e-2)
    alpha = np.linspace(0, 2 * np.pi, 20)
    ou, od = osys.scattered_field_pwe(WL, plist, ol.LayerSystem(thick, n), 1, kp, alpha)
    scale = max(np.abs(ou.coefficients).max(), np.abs(od.coefficients).max())
    for prec, tol in (('double', 1e-10), ('single', 2e-5)):
        pu, pd = scattered_field_pwe(WL, plist, LayerSystem(thick, n), 1, kp, alpha, precision=prec)
        assert np.abs(pu.coefficients - ou.coefficients).max() < tol * scale, prec
        assert np.abs(pd.coefficients - od.coefficients).max() < tol * scale, prec


def test_extinction_cross_section_matches_oracle(gpu):
    from smuthi_amd.far_field import extinction_cross_section
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.layers import LayerSystem
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[2, 1.4, 2], neff_resolution=1e-2)
    plist = [osys.Sphere([100, 100, 150], 2.4, 110, 4, 4), osys.Sphere([-100, -100, 250], 1.9, 120, 3, 3),
             osys.Sphere([-200, 100, 300], 1.7, 90, 3, 3)]
    ls = ol.LayerSystem([0, 400, 0], [2, 1.4, 2])
    for polar, pol in ((np.pi * 7 / 8, 0), (np.pi / 6, 1)):
        pw = osys.PlaneWave(WL, polar, np.pi / 3, pol)
        lsys = osys.LinearSystemOracle(plist, pw, ls, kp)
        lsys.scatter_solution(lsys.solve_lu(lsys.explicit_coupling()))
        want = osys.extinction_cross_section(pw, plist, ls)
        got = extinction_cross_section(PlaneWave(WL, polar, np.pi / 3, pol), plist, LayerSystem([0, 400, 0], [2, 1.4, 2]))
        assert abs(got - want) <= 1e-10 * abs(want)
        got32 = extinction_cross_section(PlaneWave(WL, polar, np.pi / 3, pol), plist,
                                         LayerSystem([0, 400, 0], [2, 1.4, 2]), precision='single')
        assert abs(got32 - want) <= 1e-5 * abs(want)


def test_piecewise_expansion_sums_valid_parts(gpu):
    # scattered_field_piecewise_expansion (postprocessing/scattered_field.py:13-77): layer-mediated
    # PWEs per layer plus the particles' SWEs (direct field) in their layer, nothing inside a particle;
    # points below, above and inside the particle slab
    from oracle import swe as osw
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.scattered_field import scattered_field_piecewise_expansion
    thick, n = [0, 0], [1.52, 1.0]
    plist = _particles(20, 9)
    for p in plist:
        p.position[2] = 200.0
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=n, neff_resolution=2e-2)
    alpha = np.linspace(0, 2 * np.pi, 16)
    sfld = scattered_field_piecewise_expansion(WL, plist, LayerSystem(thick, n), kp, alpha, precision='double')
    ls = ol.LayerSystem(thick, n)
    x = np.linspace(-2000, 2000, 40)
    y = np.zeros_like(x)
    k1 = ob.angular_frequency(WL) * n[1]
    swes = [(p.scattered_field, p.l_max, p.m_max, p.position, p.radius) for p in plist]
    for zz in (-300.0, 900.0, 200.0, 260.0):
        z = np.full_like(x, zz)
        z[0], x[0] = plist[3].position[2], plist[3].position[0]  # a particle centre: no field
        y0 = y.copy()
        y0[0] = plist[3].position[1]
        got = sfld.electric_field(x, y0, z, precision='double')
        layer = 0 if zz < 0 else 1
        up, dn = osys.scattered_field_pwe(WL, plist, ls, layer, kp, alpha, include_direct=False)
        parts = [dn.electric_field(x, y0, z)] if layer == 0 else [up.electric_field(x, y0, z)]
        if layer == 1:
            parts.append(osw.direct_field(swes, k1, x, y0, z, z_range=(0.0, np.inf)))
        want = [sum(p[c] for p in parts) for c in range(3)]
        inside = np.zeros(x.shape, bool)
        for p in plist:
            inside |= ~p.is_outside(x, y0, z)
        for c in range(3):
            want[c][inside] = 0
        scale = max(np.abs(w).max() for w in want)
        assert max(np.abs(g - w).max() for g, w in zip(got, want)) < 1e-10 * scale, zz
        assert all(g[0] == 0 for g in got)

def test_far_field_and_scattering_cross_section_match_oracle(gpu):
    # scattered_far_field / pwe_to_ff_conversion / scattering_cross_section
    # (postprocessing/far_field.py:223-515) on the GPU assembly vs the FP64 oracle
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd import far_field as ff
    thick, n = [0, 400, 0], [1.5, 2.0, 1.0]
    plist = _particles(30, 11)
    po = np.arange(0, np.pi + 1e-9, np.pi / 90)
    az = np.arange(0, 2 * np.pi + 1e-9, np.pi / 36)
    want = osys.scattered_far_field(WL, plist, ol.LayerSystem(thick, n), po, az)
    got = ff.scattered_far_field(WL, plist, LayerSystem(thick, n), po, az, precision='double')
    np.testing.assert_allclose(got.polar_angles, want.polar_angles, rtol=0, atol=1e-14)
    assert _rel(got.signal, want.signal) < 1e-10
    assert _rel(got.amplitude, want.amplitude) < 1e-10
    opw = osys.PlaneWave(WL, np.pi * 7 / 8, np.pi / 3, 0, 1, [0, 0, 400])
    pw = PlaneWave(WL, np.pi * 7 / 8, np.pi / 3, 0, amplitude=1, reference_point=[0, 0, 400])
    s_want = osys.total_scattering_cross_section(opw, plist, ol.LayerSystem(thick, n), po, az)
    s_got = ff.total_scattering_cross_section(initial_field=pw, particle_list=plist, layer_system=LayerSystem(thick, n),
                                              polar_angles=po, azimuthal_angles=az, precision='double')
    assert abs(s_got - s_want) / abs(s_want) < 1e-10


def test_cross_sections_of_many_solutions(gpu):
    # one particle configuration, several solution vectors: per-RHS scattering and extinction
    # cross sections equal the oracle's with each solution stored on the particles
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd import far_field as ff
    thick, n = [0, 0], [1.52, 1.0]
    plist = _particles(12, 13)
    for p in plist:
        p.position[2] = 150.0
    rng = np.random.default_rng(2)
    nu = sum(ob.blocksize(p.l_max, p.m_max) for p in plist)
    X = (rng.standard_normal((3, nu)) + 1j * rng.standard_normal((3, nu))) * 0.1
    angles = [(np.pi, 0.0, 0), (np.pi * 0.9, 0.3, 1), (np.pi * 0.8, 1.1, 0)]
    pws = [PlaneWave(WL, b, a, pol) for b, a, pol in angles]
    po = np.arange(0, np.pi + 1e-9, np.pi / 60)
    az = np.arange(0, 2 * np.pi + 1e-9, np.pi / 30)
    scs, ecs = ff.scattering_cross_sections(pws, plist, LayerSystem(thick, n), X, po, az)
    for j, (b, a, pol) in enumerate(angles):
        off = 0
        for p in plist:
            nb = ob.blocksize(p.l_max, p.m_max)
            p.scattered_field = X[j, off:off + nb]
            off += nb
        opw = osys.PlaneWave(WL, b, a, pol)
        s = osys.total_scattering_cross_section(opw, plist, ol.LayerSystem(thick, n), po, az)
        e = osys.extinction_cross_section(opw, plist, ol.LayerSystem(thick, n))
        assert abs(scs[j] - s) / abs(s) < 1e-10
        assert abs(ecs[j] - e) / abs(e) < 1e-10
