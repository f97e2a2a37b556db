"""Benchmark: RHS-solves/s of the lookup-coupling solve at 10,000 spheres and
the coupling-matvec throughput, on one MI355X.

Workload (synthetic, seeded): 10,000 spheres (radius U(80,120) nm, index
U(1.5,2.0) + 0.01i) at z = 200 nm in a 60 x 60 um square with 20 nm minimum
gaps, on glass (n = 1.52) in air, lambda = 550 nm, l_max = m_max = 3
(300,000 unknowns), radial lookup with cubic (not-a-knot spline)
interpolation at 2 nm, Sommerfeld contour with neff_resolution = neff_imag =
2e-3. A step solves a batch of right-hand sides (plane waves; incidence angle and
polarisation vary; 8 per step by default) to relative residual 1e-6: GMRES
for the whole batch in lockstep whose coupling matvecs run with FP16
operands on MFMA (one launch per 8 RHS; --precision mixed: FP32), wrapped in
FP64 iterative refinement (FP32 residuals between sweeps, the FP64 residual
certifies the result). The coefficients of right-hand side 0 are compared with a plain
FP64 GPU GMRES solve (tol 1e-9) of the same right-hand side; single-RHS
solve and plain FP64 GMRES throughputs are reported beside.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N] [--batch R]
       python bench.py --near-field    (near-field kernels at the same ensemble, one JSON line)
       python bench.py --explicit N    (explicit coupling matrix of N spheres: build and matvecs, one JSON line)
       python bench.py --rhs           (Gaussian-beam and dipole right-hand sides at the ensemble, one JSON line)
"""
import argparse
import contextlib
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MFMA_FP32_PEAK_TFLOPS = 157.3
MFMA_FP16_PEAK_TFLOPS = 2500.0
MFMA_FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
WL = 550.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload_medium():
    """Glass (n = 1.52) below air, and the Sommerfeld contour of the workload."""
    from smuthi_amd import fields as flds
    from smuthi_amd.layers import LayerSystem
    ls = LayerSystem([0, 0], [1.52, 1])
    kp = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=2e-3,
                                                 neff_imag=2e-3)
    return ls, kp


def build_workload(n_particles, seed=2):
    from smuthi_amd import ensembles
    side = 60000.0 * np.sqrt(n_particles / 10000.0)
    plist = ensembles.monolayer(n_particles, side, 200.0, seed=seed)
    ls, kp = workload_medium()
    return plist, ls, kp, side


def rhs_for_step(j):
    """Incidence of step j: polar angle pi - 8 deg * (j mod 8), polarisation j mod 2."""
    return np.pi - np.deg2rad(8.0 * (j % 8)), 0.0, j % 2


TRAFFIC_FILE = os.path.join(ROOT, 'profiles', 'radial_matvec_10k_cubic_traffic.md')
TRAFFIC_FILE16 = os.path.join(ROOT, 'profiles', 'radial_matvec_10k_fp16_8rhs.md')


def measured_traffic(kind='single'):
    """HBM bytes per matvec launch (single-RHS FP32, 8-RHS FP32, FP16 or FP64
    kernel) from the committed counter summaries, or None."""
    key = {'single': 'traffic_bytes_per_launch', 'batch': 'traffic_bytes_per_launch_8rhs',
           'batch16': 'traffic_bytes_per_launch_8rhs_fp16', 'batch64': 'traffic_bytes_per_launch_8rhs_fp64'}[kind]
    try:
        for line in open(TRAFFIC_FILE16 if kind in ('batch16', 'batch64') else TRAFFIC_FILE):
            if line.split('=')[0].strip() == key:
                return float(line.split('=')[1])
    except OSError:
        pass
    return None


def traffic_probe(n_particles, copy_bytes=512 << 20):
    """Run a plain copy of known size, then three single-RHS FP32, three 8-RHS
    FP32, three 8-RHS FP16 and three 8-RHS FP64 matvec launches of the bench
    operator (one warm-up + two timed each): the workload the FETCH_SIZE /
    WRITE_SIZE counter passes are taken on."""
    import ctypes
    from smuthi_amd import _lib, cuda
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    plist, ls, kp, side = build_workload(n_particles)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, ls, k_parallel=kp, resolution=2.0, interpolator_kind='cubic')
    ms = ctypes.c_double()
    _lib.check(_lib.load().smh_bench_copy(cuda.context(), copy_bytes, ctypes.byref(ms)), cuda.context(),
               'smh_bench_copy')
    k, _ = op.bench_matvec('single', iters=2)
    kb, _ = op.bench_matvec_batch('single', iters=2)
    kh, _ = op.bench_matvec_batch('half', iters=2)
    kd, _ = op.bench_matvec_batch('double', iters=2)
    log('copy %d bytes: %.3f ms; FP32 matvec kernel %.2f ms, 8-RHS FP32 kernel %.2f ms, 8-RHS FP16 kernel %.2f ms, '
        '8-RHS FP64 kernel %.2f ms' % (copy_bytes, ms.value, k, kb, kh, kd))


# flops per (point, particle) pair of the SWE direct sum at l_max = m_max = 3: 38 per (l, m) for both
# tau, 50 per m (rotation and Cartesian accumulation), ~150 for geometry, Hankel and Legendre
SWE_FLOPS_PER_PAIR = 15 * 38 + 7 * 50 + 150
# flop-equivalents per (point, kappa, alpha) term of the PWE field sum (3 complex FMAs, phase,
# 3 transcendentals at quarter rate)
PWE_FLOP_EQ_PER_TERM = 62


def near_field_bench(n_particles, n_grid=512, n_kappa=2000, n_alpha=180, seed=7):
    """Time the two near-field kernels at the scale of the bench ensemble:
    the direct field of every particle's outgoing SWE (l_max = m_max = 3) and
    one plane wave expansion (n_kappa x n_alpha), both on an n_grid^2 xz-plane
    grid above the particle slab (x in [-30, 30] um, z in [330, 3000] nm).
    Expansion coefficients are synthetic (seeded); the kernel time does not
    depend on their values. HIP-event times of the kernels alone."""
    from smuthi_amd import cuda, fields as flds
    from smuthi_amd.expansions import PlaneWaveExpansion, swe_field_sum
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    plist, ls, kp_, side = build_workload(n_particles)
    rng = np.random.default_rng(seed)
    k = 2 * np.pi / WL
    B = flds.blocksize(3, 3)
    pos = np.array([p.position for p in plist], dtype=float)
    rad = np.array([p.radius for p in plist], dtype=float)
    zlim = np.tile([0.0, np.inf], (len(plist), 1))
    coef = (rng.standard_normal((len(plist), B)) + 1j * rng.standard_normal((len(plist), B))) * 1e-3
    X, Z = np.meshgrid(np.linspace(-side / 2, side / 2, n_grid), np.linspace(330.0, 3000.0, n_grid))
    pts = np.stack([X.ravel(), np.zeros(X.size), Z.ravel()], axis=1)
    pairs = float(len(pts)) * len(plist)
    out = {'metric': 'near-field kernels, %d spheres, %dx%d points' % (len(plist), n_grid, n_grid),
           'points': len(pts), 'particles': len(plist), 'data': 'synthetic expansion coefficients'}
    for prec, peak in (('single', 157.3), ('double', MFMA_FP64_PEAK_TFLOPS)):
        swe_field_sum(pos, rad, zlim, coef, 3, 3, k, pts[:4096], 0, 1.0, prec)  # warm-up
        swe_field_sum(pos, rad, zlim, coef, 3, 3, k, pts, 0, 1.0, prec)
        ms = swe_field_sum.last_kernel_ms
        tf = pairs * SWE_FLOPS_PER_PAIR / (ms * 1e-3) / 1e12
        out['swe_direct_sum_' + ('fp32' if prec == 'single' else 'fp64')] = {
            'kernel': 'swe_field_kernel<%s,3,3>' % ('float' if prec == 'single' else 'double'), 'ms': ms,
            'pairs': pairs, 'svwf_terms': pairs * B, 'flops_per_pair': SWE_FLOPS_PER_PAIR, 'tflops': tf,
            'peak': peak, 'frac': tf / peak}
        log('SWE direct sum %s: %.2f ms, %.1f TFLOP/s' % (prec, ms, tf))
    kpw = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1.4e-3,
                                                  neff_imag=1e-2)
    kpw = kpw[:n_kappa] if len(kpw) >= n_kappa else np.concatenate(
        [kpw, kpw[-1] + np.arange(1, n_kappa + 1 - len(kpw)) * (kpw[-1] - kpw[-2])])
    alpha = np.linspace(0, 2 * np.pi, n_alpha)
    pwe = PlaneWaveExpansion(k, kpw, alpha, 'upgoing', [0, 0, 0], 0, np.inf)
    pwe.coefficients = (rng.standard_normal((2, n_kappa, n_alpha)) + 1j * rng.standard_normal((2, n_kappa, n_alpha)))
    pwe.electric_field(pts[:4096, 0], pts[:4096, 1], pts[:4096, 2], precision='single')  # warm-up
    pwe.electric_field(pts[:, 0], pts[:, 1], pts[:, 2], precision='single')
    ms = pwe.last_kernel_ms
    terms = float(len(pts)) * n_kappa * n_alpha
    tf = terms * PWE_FLOP_EQ_PER_TERM / (ms * 1e-3) / 1e12
    out['pwe_field_fp32'] = {'kernel': 'pwe_field_kernel<float>', 'ms': ms, 'terms': terms,
                             'flop_eq_per_term': PWE_FLOP_EQ_PER_TERM, 'tflops_equiv': tf, 'peak': 157.3,
                             'frac': tf / 157.3}
    log('PWE field fp32: %.2f ms, %.1f TFLOP/s-equivalent' % (ms, tf))
    # scattered far field of one solution vector on the default half-degree grid
    from smuthi_amd import far_field as ffm
    from smuthi_amd.scattered_field import scattered_field_pwe
    az, po = flds.angular_arrays(np.pi / 360)
    x = (rng.standard_normal(len(plist) * B) + 1j * rng.standard_normal(len(plist) * B)) * 1e-3
    parts = ffm._per_particle(plist, x)
    for prec in ('single', 'double'):
        t0 = time.time()
        ffm.scattered_far_field(WL, parts, ls, po, az, precision=prec)
        wall = time.time() - t0
        kms = scattered_field_pwe.last_kernel_ms  # the bottom-hemisphere call
        t0 = time.time()
        ffm.scattered_far_field(WL, parts, ls, po, az, precision=prec)
        wall = time.time() - t0
        out['scattered_far_field_' + ('fp32' if prec == 'single' else 'fp64')] = {
            'seconds': wall, 'polar': len(po), 'azimuthal': len(az),
            'assembly_kernel_ms_bottom': kms,
            'terms_per_hemisphere': float(len(plist)) * (len(po) // 2) * len(az)}
        log('scattered far field %s: %.3f s' % (prec, wall))
    return json.dumps(out)


# executed FP64 flops per (receiver, emitter, kappa, matrix entry) of the
# explicit layer-mediated sum: t = d0 T1 + d1 T0 (2 complex products and a
# sum, 14 flops) and acc += J t (a complex multiply-add, 8 flops)
EXPLICIT_FLOPS_PER_TERM = 22


def explicit_bench(n_particles=1000, seed=11, cpu_pairs=20):
    """Build the explicit (dense FP64) coupling matrix of n_particles spheres
    at heights U(200, 400) nm in air above glass (lambda 550 nm, l_max = m_max
    = 3, the bench's Sommerfeld contour), then time its dense matvecs.
    Reports the build kernels' HIP-event time, the layer-mediated sum's FP64
    rate against the FP64 peak, the matvecs' W bandwidth against HBM, and the
    oracle's per-pair CPU time (explicit direct + layer-mediated blocks, as
    smuthi's CouplingMatrixExplicit builds them) on a sample of pairs."""
    from smuthi_amd import cuda, ensembles, fields as flds
    from smuthi_amd.linear_system import CouplingMatrixExplicit
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    ls, kp = workload_medium()
    side = 60000.0 * np.sqrt(n_particles / 10000.0)
    flat = ensembles.monolayer(n_particles, side, 300.0, seed=seed)
    rng = np.random.default_rng(seed)
    for p in flat:
        p.position = [p.position[0], p.position[1], float(rng.uniform(200.0, 400.0))]
    B = flds.blocksize(3, 3)
    n = n_particles * B
    log('explicit: %d spheres, %d unknowns, N_kappa %d' % (n_particles, n, len(kp)))
    CouplingMatrixExplicit(WL, flat[:64], ls, k_parallel=kp)  # warm-up (kernel load)
    t0 = time.time()
    op = CouplingMatrixExplicit(WL, flat, ls, k_parallel=kp)
    wall = time.time() - t0
    terms = float(n_particles) ** 2 * len(kp) * B * B
    tf = terms * EXPLICIT_FLOPS_PER_TERM / (op.build_kernel_ms * 1e-3) / 1e12
    out = {'metric': 'explicit coupling matrix build, %d spheres (%d unknowns)' % (n_particles, n),
           'value': op.build_kernel_ms / 1e3, 'unit': 's (build kernels)', 'higher_is_better': False,
           'wall_s': wall, 'host_coefficients_s': op.host_seconds, 'n_kappa': len(kp), 'dtype': 'f64',
           'data': 'synthetic (seeded positions)',
           'roofline': {'bound': 'fp64', 'achieved': tf, 'peak': MFMA_FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                        'frac': tf / MFMA_FP64_PEAK_TFLOPS, 'traffic': None,
                        'note': 'all build kernels timed together; %d flops per (pair, kappa, entry)'
                                % EXPLICIT_FLOPS_PER_TERM}}
    log('explicit build: %.3f s kernels (%.1f TFLOP/s), %.2f s wall' % (op.build_kernel_ms / 1e3, tf, wall))
    for prec, esz in (('double', 16), ('single', 8)):
        k_ms, t_ms = op.bench_matvec(prec, iters=10)
        gbs = float(n) * n * esz / (k_ms * 1e-3) / 1e9
        out['matvec_' + ('fp64' if prec == 'double' else 'fp32')] = {
            'kernel': 'explicit_gemv_kernel<%s>' % ('double' if prec == 'double' else 'float'), 'ms': k_ms,
            'matvec_ms': t_ms, 'GB/s': gbs, 'frac_hbm': gbs / HBM_PEAK_GBS}
        log('explicit matvec %s: %.3f ms, %.0f GB/s' % (prec, k_ms, gbs))
    # CPU: the oracle's explicit blocks (smuthi's per-pair Python path) on a sample of pairs
    from oracle import coupling as oc, layers as ol
    from oracle.system import Sphere as OSphere
    ls_o = ol.LayerSystem(ls.thicknesses, ls.refractive_indices)
    sample = [OSphere(p.position, p.refractive_index, p.radius, 3, 3) for p in flat[:cpu_pairs + 1]]
    t0 = time.time()
    for q in range(cpu_pairs):
        oc.direct_coupling_block(WL, sample[0], sample[q + 1], ls_o)
        oc.layer_mediated_coupling_block(WL, sample[0], sample[q + 1], ls_o, kp)
    per_pair = (time.time() - t0) / cpu_pairs
    out['cpu_baseline'] = {'value': per_pair * n_particles ** 2, 'unit': 's (extrapolated build)', 'cores': 1,
                           'kind': 'port', 'sample': '%d pairs, %.3f s per pair' % (cpu_pairs, per_pair)}
    log('oracle CPU: %.3f s per pair -> %.0f s for the matrix' % (per_pair, per_pair * n_particles ** 2))
    return json.dumps(out)


# FP64 flop-equivalents per (point, kappa, alpha) term of the PWE-to-SWE
# kernel: u = cos(a) dx + sin(a) dy (3), e^{i kappa u} (sincos and, off the
# real axis, exp: counted as 20), 2 (2 m_max + 1) complex multiply-adds (8 each)
P2S_FLOP_EQ_PER_TERM_33 = 3 + 20 + 14 * 8


def rhs_bench(n_particles, n_dipoles=16, seed=13, cpu_particles=3):
    """Right-hand sides other than plane waves at the bench ensemble: the
    regular SWE coefficients of a Gaussian beam (waist 5 um, from the top, on
    the bench's Sommerfeld contour truncated to the top medium's propagating
    range x the default 0.5-degree azimuthal grid) and of a collection of
    point dipoles (explicit blocks from every dipole to every sphere). HIP-event
    kernel times; the oracle's per-particle CPU time beside."""
    from smuthi_amd import cuda, fields as flds
    from smuthi_amd import initial_field as inf
    from smuthi_amd import linear_system as lsm
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    plist, ls, kp, side = build_workload(n_particles)
    k0 = 2 * np.pi / WL
    kpb = kp[np.real(kp) < 0.999 * k0]
    al, _ = flds.angular_arrays(np.pi / 360)
    beam = inf.GaussianBeam(WL, np.pi, 0.0, 0, 5000.0, k_parallel_array=kpb, azimuthal_angles_array=al)
    beam.swe_coefficients(plist[:256], ls)  # warm-up
    t0 = time.time()
    beam.swe_coefficients(plist, ls)
    wall = time.time() - t0
    ms = inf.pwe_to_swe.last_kernel_ms  # the downgoing expansion's launch (the upgoing one is the same size)
    terms = float(len(plist)) * len(kpb) * len(al)
    tf = terms * P2S_FLOP_EQ_PER_TERM_33 / (ms * 1e-3) / 1e12
    out = {'metric': 'non-plane-wave right-hand sides, %d spheres' % len(plist), 'dtype': 'f64',
           'data': 'synthetic (seeded positions)',
           'gaussian_beam': {'kernel': 'pwe_to_swe_kernel<3,3>', 'ms_per_expansion': ms, 'wall_s_both': wall,
                             'n_kappa': len(kpb), 'n_alpha': len(al), 'terms': terms,
                             'flop_eq_per_term': P2S_FLOP_EQ_PER_TERM_33, 'tflops_equiv': tf,
                             'peak': MFMA_FP64_PEAK_TFLOPS, 'frac': tf / MFMA_FP64_PEAK_TFLOPS}}
    log('gaussian beam: %.2f ms per expansion (%.1f TFLOP/s-eq), %.2f s wall' % (ms, tf, wall))
    rng = np.random.default_rng(seed)
    coll = inf.DipoleCollection(WL, k_parallel_array=kp)
    for _ in range(n_dipoles):
        pos = [rng.uniform(-side / 2, side / 2), rng.uniform(-side / 2, side / 2), 450.0]
        coll.append(inf.DipoleSource(WL, list(rng.standard_normal(3) * 1e7), pos, k_parallel_array=kp))
    coll.swe_coefficients(plist[:256], ls)  # warm-up
    t0 = time.time()
    coll.swe_coefficients(plist, ls)
    wall = time.time() - t0
    kms = lsm.explicit_blocks.last_kernel_ms
    layer_flops = float(len(plist)) * n_dipoles * len(kp) * 900 * EXPLICIT_FLOPS_PER_TERM
    out['dipoles'] = {'dipoles': n_dipoles, 'kernel_ms': kms, 'wall_s': wall,
                      'tflops': layer_flops / (kms * 1e-3) / 1e12, 'n_kappa': len(kp)}
    log('%d dipoles: %.2f ms kernels, %.2f s wall' % (n_dipoles, kms, wall))
    # CPU: the oracle's Gaussian-beam conversion per particle
    from oracle import initial as oin, layers as ol, system as osys
    ls_o = ol.LayerSystem(ls.thicknesses, ls.refractive_indices)
    beam_o = oin.GaussianBeam(WL, np.pi, 0.0, 0, 5000.0, kpb, al)
    t0 = time.time()
    for p in plist[:cpu_particles]:
        osys.initial_field_coefficients(beam_o, osys.Sphere(p.position, 1.5, 100, 3, 3), ls_o)
    per = (time.time() - t0) / cpu_particles
    out['cpu_baseline'] = {'value': per * len(plist), 'unit': 's (extrapolated, Gaussian beam)', 'cores': 1,
                           'kind': 'port', 'sample': '%d particles, %.2f s each' % (cpu_particles, per)}
    log('oracle CPU: %.2f s per particle' % per)
    return json.dumps(out)


def cpu_baseline(table, rho, n_full, iters_per_solve, n_sample=400, seed=5, budget_s=20.0):
    """Time the oracle's restatement of smuthi's CPU radial-lookup matvec
    (CouplingMatrixRadialLookupCPU.matvec: per-(n1, n2) scipy interp1d 'cubic',
    gather, phase, dot) on a sample ensemble of the same density, scale the
    matvec by (N / N_sample)^2 and multiply by the GMRES iteration count."""
    from oracle import coupling as oc
    from oracle.system import Sphere
    from smuthi_amd import ensembles
    side = 60000.0 * np.sqrt(n_sample / 10000.0)
    pp = ensembles.monolayer(n_sample, side, 200.0, seed=seed)
    plist = [Sphere(p.position, p.refractive_index, p.radius, 3, 3) for p in pp]
    # the table of the full ensemble covers every distance of the sample
    loop = oc.RadialLookupCPUReferenceLoop(table, rho, plist, 'cubic')
    n_unk = 30 * n_sample
    v = np.random.default_rng(1).standard_normal(n_unk) + 0j
    rows, t_used = [], 0.0
    t0 = time.time()
    for n1 in range(30):
        loop.matvec(v, n1_subset=[n1])
        rows.append(n1)
        t_used = time.time() - t0
        if t_used > budget_s:
            break
    t_matvec_sample = t_used * 30 / len(rows)
    t_matvec_full = t_matvec_sample * (n_full / n_sample) ** 2
    t_solve = t_matvec_full * max(iters_per_solve, 1)
    return dict(value=1.0 / t_solve, unit='RHS-solves/s', cores=1, kind='port',
                sample=('smuthi CPU radial-lookup matvec restated (oracle RadialLookupCPUReferenceLoop, scipy '
                        'interp1d cubic), %d of 30 row groups of a %d-sphere sample at the same density, %.1f s; '
                        'scaled by (N/N_sample)^2 to %.1f s per matvec and by %d GMRES iterations per solve'
                        % (len(rows), n_sample, t_used, t_matvec_full, iters_per_solve)),
                matvec_seconds_est=t_matvec_full)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=2)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--particles', type=int, default=10000)
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--batch', type=int, default=8, help='right-hand sides solved together per step')
    ap.add_argument('--precision', default='mixed16', choices=['mixed16', 'mixed', 'double'],
                    help='inner matvec precision of the timed solves (mixed16: FP16 operands, mixed: FP32)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--matvec-iters', type=int, default=5)
    ap.add_argument('--traffic-probe', action='store_true',
                    help='only run the copy calibration and three FP32 matvecs (for counter passes)')
    ap.add_argument('--explicit', type=int, default=0, metavar='N',
                    help='build the explicit coupling matrix of N spheres and time its matvecs (one JSON line)')
    ap.add_argument('--rhs', action='store_true',
                    help='time the Gaussian-beam and dipole right-hand sides at the ensemble (one JSON line)')
    ap.add_argument('--near-field', action='store_true',
                    help='time the SWE direct-sum and PWE near-field kernels instead (one JSON line)')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)', file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe(args.particles)
        return None
    if args.near_field:
        return near_field_bench(args.particles)
    if args.explicit:
        return explicit_bench(args.explicit)
    if args.rhs:
        return rhs_bench(args.particles)

    from smuthi_amd import cuda
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.linear_system import LinearSystem
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    dev, n_cu = cuda.device_info()

    plist, ls, kp, side = build_workload(args.particles)
    log('workload: %d spheres, side %.0f nm, N_kappa %d' % (len(plist), side, len(kp)))

    def make_system(j, precision, tol):
        polar, az, pol = rhs_for_step(j)
        return LinearSystem(plist, PlaneWave(WL, polar, az, pol), ls, kp, solver_type='gmres', solver_tolerance=tol,
                            store_coupling_matrix=False, coupling_matrix_lookup_resolution=2.0, interpolator_kind='cubic',
                            precision=precision)

    t0 = time.time()
    base = make_system(0, 'mixed', args.tol)
    base.prepare()
    prep_s = time.time() - t0
    op = base.coupling_matrix
    log('prepare %.2f s (table %.2f s), len_rho %d' % (prep_s, op.table_seconds, len(op.radial_distance_array)))

    # right-hand sides of all steps (setup, outside the timed region): a step is a
    # batch of `batch` incidence directions / polarisations
    nb = args.batch
    # (T is the same for every incidence: rhs_j = T a_j)
    rhs = []
    for j in range(nb * (args.warmup + args.steps)):
        polar, az, pol = rhs_for_step(j)
        rhs.append(base.t_matrix.diagonal * PlaneWave(WL, polar, az, pol).swe_coefficients(plist, ls))
    rhs = np.stack(rhs)

    # coupling matvec kernels (HIP events on the operator's stream)
    k32, tot32 = op.bench_matvec('single', iters=args.matvec_iters)
    k64, tot64 = op.bench_matvec('double', iters=max(2, args.matvec_iters // 2))
    kb32, totb32 = op.bench_matvec_batch('single', iters=max(2, args.matvec_iters // 2))
    kb16, totb16 = op.bench_matvec_batch('half', iters=max(2, args.matvec_iters // 2))
    kb64, totb64 = op.bench_matvec_batch('double', iters=2)
    n_unk = op.shape[0]
    E = float(n_unk) ** 2
    flops_1 = E * (4 * 4 + 8 * 1)  # cubic: 4 taps per complex entry, one RHS
    flops_8 = E * (4 * 4 + 8 * 8)  # the same entries applied to 8 RHS
    # the FP16 inner operator interpolates with its own number of Horner terms (2: linear between nodes)
    half_terms, half_lin_ratio = op.half_interpolation()
    flops_16 = E * (4 * half_terms + 8 * 8)
    log('matvec FP32 %.2f ms, FP64 %.2f ms; 8-RHS FP16 %.2f ms, FP32 %.2f ms, FP64 %.2f ms'
        % (k32, k64, kb16, kb32, kb64))

    def solve(j0, count, precision, tol):
        sysj = make_system(j0, precision, tol)
        sysj.t_matrix = base.t_matrix
        sysj.coupling_matrix = op
        x = sysj.solve(rhs[j0:j0 + count])
        return x, dict(sysj.solve_stats)

    # FP64 reference for right-hand side 0 and the plain FP64 GMRES throughput (one RHS)
    t0 = time.time()
    x_ref = solve(0, 1, 'double', 1e-9)[0][0]
    log('FP64 reference solve (tol 1e-9): %.2f s' % (time.time() - t0))
    t0 = time.time()
    _, plain_stats = solve(0, 1, 'double', args.tol)
    fp64_plain_s = time.time() - t0
    log('plain FP64 GMRES (tol %.0e): %.2f s, %d iterations' % (args.tol, fp64_plain_s, plain_stats['iterations']))
    # one mixed-precision solve of a single RHS (no batching) for comparison
    t0 = time.time()
    x_single, single_stats = solve(0, 1, 'mixed', args.tol)
    single_s = time.time() - t0
    log('single-RHS mixed solve: %.2f s' % single_s)

    # timed region: batched solves at the requested inner precision
    prec = args.precision
    for w in range(args.warmup):
        solve(nb * w, nb, prec, args.tol)
    t0 = time.time()
    sols, stats = [], []
    for st in range(args.warmup, args.warmup + args.steps):
        x, stt = solve(nb * st, nb, prec, args.tol)
        sols.append(x)
        stats.append(stt)
    elapsed = time.time() - t0
    # accuracy: RHS 0 solved inside a batch vs the FP64 solution
    x_b0 = solve(0, nb, prec, args.tol)[0][0]
    coef_err = float(np.linalg.norm(x_b0 - x_ref) / np.linalg.norm(x_ref))
    single_err = float(np.linalg.norm(x_single[0] - x_ref) / np.linalg.norm(x_ref))
    worst_res = max(s['relative_residual'] for s in stats)
    ok = coef_err <= 1e-5 and worst_res <= args.tol
    n_solved = nb * args.steps
    value = n_solved / elapsed if ok else 0.0
    log('batched %s solves: %.2f s for %d RHS, residual %.2e, coefficient error vs FP64 %.2e'
        % (prec, elapsed, n_solved, worst_res, coef_err))

    cpu = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(op.lookup_table, op.radial_distance_array, len(plist), plain_stats['iterations'])
        log('cpu baseline %.1f s' % (time.time() - t0))

    batched = nb > 1
    half = prec == 'mixed16'
    kern_ms = (kb16 if half else kb32) if batched else k32
    flops_k = ((flops_16 if half else flops_8) if batched else flops_1)
    achieved = flops_k / (kern_ms * 1e-3) / 1e12
    peak = MFMA_FP16_PEAK_TFLOPS if (half and batched) else MFMA_FP32_PEAK_TFLOPS
    out = {
        'metric': 'RHS-solves/s at 10k spheres, err<=1e-5 vs FP64; coupling matvec GB/s',
        'value': value,
        'unit': 'RHS-solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': elapsed / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': ('fp16 operands / fp32 accumulation coupling matvec (MFMA), fp32 sweep residuals, '
                  'fp64 refinement/Arnoldi/certifying residual' if half else
                  'fp32 coupling matvec (MFMA), fp64 refinement/Arnoldi/residual'),
        'data': 'synthetic',
        'config': {
            'workload': ('radial-lookup coupling solve, %d spheres monolayer on glass, l_max=3, cubic spline 2 nm, '
                         'GMRES tol %.0e, %s-in-FP64 refinement, %d right-hand sides (plane waves) per step'
                         % (len(plist), args.tol, 'FP16' if half else ('FP32' if prec == 'mixed' else 'no'), nb)),
            'inner_precision': prec,
            'particles': len(plist), 'unknowns': n_unk, 'lookup_rows': int(len(op.radial_distance_array)),
            'n_kappa': int(len(kp)), 'side_nm': side, 'seed': 2, 'tolerance': args.tol,
            'interpolation': 'cubic (not-a-knot spline)', 'resolution_nm': 2.0, 'rhs_per_step': nb,
        },
        'coupling_matvec_gbps_equiv_dense': E * 8 / (k32 * 1e-3) / 1e9,
        'matvec_ms': {'fp32_kernel': k32, 'fp32_total': tot32, 'fp64_kernel': k64, 'fp64_total': tot64,
                      'fp16_kernel_8rhs': kb16, 'fp16_total_8rhs': totb16,
                      'fp32_kernel_8rhs': kb32, 'fp32_total_8rhs': totb32, 'fp64_kernel_8rhs': kb64,
                      'fp64_total_8rhs': totb64},
        'coef_err_vs_fp64': coef_err,
        'worst_relative_residual': worst_res,
        'solve_stats': stats,
        'single_rhs_mixed': {'seconds': single_s, 'rhs_solves_per_s': 1.0 / single_s, 'coef_err_vs_fp64': single_err,
                             **single_stats},
        'fp64_plain_gmres': {'seconds': fp64_plain_s, 'rhs_solves_per_s': 1.0 / fp64_plain_s, **plain_stats},
        'speedup_vs_fp64_gmres': (value * fp64_plain_s) if value else None,
        'prepare_seconds': prep_s,
        'table_seconds': op.table_seconds,
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s',
                     'frac': achieved / peak,
                     'traffic': measured_traffic(('batch16' if half else 'batch') if batched else 'single'),
                     'kernel': (('batch_half_kernel<3,3,%d>' % half_terms if half else 'batch_slots_kernel<3,3,4>')
                                if batched else 'lookup_matvec_kernel<float,float,3,3,4,false>'),
                     'flops_per_launch': flops_k},
        'fp16_inner_operator': {'interpolation': ('linear between table nodes' if half_terms == 2
                                                  else 'the operator\'s own (%d Horner terms)' % half_terms),
                                'linear_error_over_bin_scale': half_lin_ratio},
        'roofline_fp32_8rhs_kernel': {'kernel': 'batch_slots_kernel<3,3,4>', 'ms': kb32,
                                      'achieved': flops_8 / (kb32 * 1e-3) / 1e12, 'peak': MFMA_FP32_PEAK_TFLOPS,
                                      'frac': flops_8 / (kb32 * 1e-3) / 1e12 / MFMA_FP32_PEAK_TFLOPS,
                                      'traffic': measured_traffic('batch')},
        'roofline_fp64_8rhs_kernel': {'kernel': 'batch_slots64_kernel<3,3,4>', 'ms': kb64,
                                      'achieved': flops_8 / (kb64 * 1e-3) / 1e12, 'peak': MFMA_FP64_PEAK_TFLOPS,
                                      'frac': flops_8 / (kb64 * 1e-3) / 1e12 / MFMA_FP64_PEAK_TFLOPS,
                                      'traffic': measured_traffic('batch64')},
        'roofline_single_rhs_kernel': {'achieved': flops_1 / (k32 * 1e-3) / 1e12, 'peak': MFMA_FP32_PEAK_TFLOPS,
                                       'frac': flops_1 / (k32 * 1e-3) / 1e12 / MFMA_FP32_PEAK_TFLOPS,
                                       'traffic': measured_traffic('single')},
        'cpu_baseline': cpu,
        'device': dev, 'compute_units': n_cu, 'host': platform.node(), 'host_cpus': os.cpu_count(),
    }
    return json.dumps(out)


if __name__ == '__main__':
    # the drop-in LinearSystem reports 'Number of unknowns' on stdout like smuthi does;
    # keep stdout to the one JSON line
    _stdout = sys.stdout
    with contextlib.redirect_stdout(sys.stderr):
        _line = main()
    if _line is not None:
        print(_line, file=_stdout, flush=True)
