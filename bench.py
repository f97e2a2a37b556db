"""Benchmark: RHS-solves/s of the lookup-coupling solve at 10,000 spheres and
the coupling-matvec throughput, on one MI355X.

Workload (synthetic, seeded): 10,000 spheres (radius U(80,120) nm, index
U(1.5,2.0) + 0.01i) at z = 200 nm in a 60 x 60 um square with 20 nm minimum
gaps, on glass (n = 1.52) in air, lambda = 550 nm, l_max = m_max = 3
(300,000 unknowns), radial lookup with cubic (not-a-knot spline)
interpolation at 2 nm, Sommerfeld contour with neff_resolution = neff_imag =
2e-3. The right-hand sides are 64 distinct plane waves (32 polar angles from
0 to 60 degrees, TE and TM); a step solves the next 8 of them to relative
residual 1e-6: GMRES for the whole batch in lockstep whose coupling matvecs
run with FP16 operands on MFMA (one launch per 8 RHS; --precision mixed:
FP32), wrapped in FP64 iterative refinement (FP32 residuals between sweeps,
the FP64 residual certifies the result). Every timed right-hand side is
compared with an FP64 GMRES solve (tol 1e-10) of it, and only those within
1e-5 count. Batched FP64 GMRES, single-RHS FP64 GMRES and the FP32-inner
refinement are timed on the same incidences beside it, the CPU baseline on
the host's cores. Before timing, the autotuner picks the matvec kernels'
launch shapes (setup, reported under "autotune").

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--particles N] [--batch R]
       python bench.py --config 3   (coefficient, extinction and |E| errors of the FP32 pipeline vs FP64, l_max 3 and 4)
       python bench.py --config 4   (ensembles of uneven size sharing one volume table, one JSON line)
       python bench.py --config 5   (learned near-field preconditioner vs plain GMRES, one JSON line)
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
    operator (one warm-up + two timed each) after the headline's autotuning:
    the workload the FETCH_SIZE / WRITE_SIZE counter passes are taken on."""
    import ctypes
    from smuthi_amd import _lib, cuda
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    plist, ls, kp, side = build_workload(n_particles)
    op = CouplingMatrixRadialLookupCUDA(WL, plist, ls, k_parallel=kp, resolution=2.0, interpolator_kind='cubic')
    # the launch shapes of the headline run (its autotuner's choices), so the counters describe those kernels;
    # the counted launches are the last three of each kernel
    autotune(op, ((True, 'half'), (True, 'double'), (False, 'single')))
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


def incidences():
    """The 64 right-hand sides of the batching configuration: polar angles pi - theta_j, theta_j = j 60 deg / 31
    (j = 0..31), azimuth 0, TE and TM."""
    th = np.deg2rad(np.arange(32) * 60.0 / 31)
    return [(np.pi - t, 0.0, pol) for t in th for pol in (0, 1)]


def _child(args, timeout=600):
    """Run `python -m oracle.cpu_bench ...` (a child process that never touches the GPU) and parse its line."""
    import subprocess
    res = subprocess.run([sys.executable, '-m', 'oracle.cpu_bench'] + [str(a) for a in args], cwd=ROOT,
                         capture_output=True, text=True, timeout=timeout)
    if res.returncode != 0:
        raise RuntimeError('CPU baseline failed: %s' % res.stderr[-2000:])
    return json.loads(res.stdout.strip().splitlines()[-1])


def cpu_baseline(table, rho, n_full, iters_per_solve, workdir, n_sample=400, seed=5):
    """The oracle's restatement of smuthi's CPU radial-lookup matvec (CouplingMatrixRadialLookupCPU.matvec:
    per-(n1, n2) scipy interp1d 'cubic', gather, phase, dot) on a sample ensemble of the same density, its 30
    row groups over all host cores (up to 16, one BLAS thread each); the matvec time is scaled by
    (N / N_sample)^2 and multiplied by the FP64 GMRES iteration count."""
    from smuthi_amd import ensembles
    side = 60000.0 * np.sqrt(n_sample / 10000.0)
    pp = ensembles.monolayer(n_sample, side, 200.0, seed=seed)
    pos = np.array([p.position for p in pp])
    d = pos[:, None, :2] - pos[None, :, :2]
    rmax = float(np.sqrt((d ** 2).sum(-1)).max())
    rows = int(np.searchsorted(rho, rmax + 3 * (rho[1] - rho[0]))) + 3
    paths = [os.path.join(workdir, f) for f in ('cpu_table.npy', 'cpu_rho.npy', 'cpu_pos.npy')]
    np.save(paths[0], np.ascontiguousarray(table[:rows]))
    np.save(paths[1], rho[:rows])
    np.save(paths[2], pos)
    workers = max(1, min(16, os.cpu_count() or 1))
    r = _child(['matvec'] + paths + ['cubic', workers])
    t_full = r['matvec_seconds'] * (n_full / n_sample) ** 2
    t_solve = t_full * max(iters_per_solve, 1)
    return dict(value=1.0 / t_solve, unit='RHS-solves/s', cores=r['workers'], kind='port',
                sample=('smuthi CPU radial-lookup matvec restated (oracle RadialLookupCPUReferenceLoop, scipy interp1d '
                        'cubic) on a %d-sphere sample at the same density: all 30 row groups over %d worker processes '
                        '(1 BLAS thread each; host has %s CPUs, default BLAS threads %s) in %.2f s; scaled by '
                        '(N/N_sample)^2 to %.1f s per matvec and by %d FP64 GMRES iterations per solve'
                        % (n_sample, r['workers'], r['cpu_count'], r['blas_threads_default'], r['matvec_seconds'],
                           t_full, iters_per_solve)),
                matvec_seconds_est=t_full, host_cpus=r['cpu_count'])


def cfg1_cpu_vs_gpu(workdir, tol=1e-5):
    """The CPU-runnable configuration (50 spheres r = 100 nm, n = 1.5 + 0.01i at z = 200 nm on glass, radial
    linear 5 nm, one plane wave, GMRES tol 1e-5): smuthi's CPU lookup solve restated (oracle matvec in scipy
    GMRES, one core) beside the GPU drop-in (FP64 GMRES), same table; solutions compared."""
    from smuthi_amd import ensembles, fields as flds
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import LinearSystem
    plist = ensembles.small_substrate_ensemble(seed=1)
    ls = LayerSystem([0, 0], [1.52, 1])
    kp = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    lsys = LinearSystem(plist, PlaneWave(WL, np.pi, 0.0, 0), ls, kp, solver_type='gmres', solver_tolerance=tol,
                        store_coupling_matrix=False, coupling_matrix_lookup_resolution=5.0, interpolator_kind='linear',
                        precision='double')
    lsys.prepare()
    t0 = time.time()
    x_gpu = lsys.solve()
    gpu_s = time.time() - t0
    op = lsys.coupling_matrix
    paths = [os.path.join(workdir, f) for f in ('cfg1_table.npy', 'cfg1_rho.npy', 'cfg1.npz')]
    np.save(paths[0], op.lookup_table)
    np.save(paths[1], op.radial_distance_array)
    np.savez(paths[2], pos=np.array([p.position for p in plist]), n=np.array([p.refractive_index for p in plist]),
             r=np.array([p.radius for p in plist]), t=lsys.t_matrix.diagonal, rhs=lsys.t_matrix.right_hand_side())
    r = _child(['solve'] + paths + ['linear', tol])
    x_cpu = np.load(paths[2].replace('.npz', '_x.npy'))
    return dict(cpu_seconds=r['seconds'], cpu_iterations=r['iterations'], cpu_relative_residual=r['relative_residual'],
                cpu_rhs_solves_per_s=1.0 / r['seconds'], cpu_cores=1, gpu_seconds=gpu_s,
                gpu_rhs_solves_per_s=1.0 / gpu_s, gpu_iterations=lsys.solve_stats['iterations'],
                gpu_relative_residual=lsys.solve_stats['relative_residual'],
                solution_difference=float(np.linalg.norm(x_gpu - x_cpu) / np.linalg.norm(x_cpu)),
                workload='50 spheres on glass, radial linear 5 nm, FP64 GMRES tol %.0e (scipy on the CPU, GPU '
                         'drop-in), one plane wave' % tol)


def _order_name(order):
    return {0: 'chunk', 1: 'xcd'}.get(order, 'bands%d' % order)


def autotune(op, shapes, enabled=True):
    """Launch-shape autotuning of the operator's kernels (setup, outside the timed region): per shape
    (batch, precision) the kernel time of every candidate grouping and the one kept."""
    if not enabled:
        return None
    out = {}
    t0 = time.time()
    for batch, prec in shapes:
        r = op.autotune(batch=batch, precision=prec, iters=2)
        out['%s_%s' % ('8rhs' if batch else '1rhs', prec)] = {
            'best': {'items_per_cu': r['best'][0], 'launch_order': _order_name(r['best'][1])},
            'matvec_ms': {'%d/%s' % (k[0], _order_name(k[1])): round(v, 4) for k, v in r['matvec_ms'].items()}}
        log('autotune %s %s: best %s, %.3f ms (default 16/xcd %.3f ms)'
            % ('8-RHS' if batch else '1-RHS', prec, r['best'], r['matvec_ms'][r['best']],
               r['matvec_ms'].get((16, 1), float('nan'))))
    out['seconds'] = time.time() - t0
    return out


def _errors(x, ref):
    return [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in zip(x, ref)]


def main_headline(args):
    import tempfile
    from smuthi_amd import cuda
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.linear_system import LinearSystem
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    dev, n_cu = cuda.device_info()
    workdir = tempfile.mkdtemp(prefix='smh_bench_')

    plist, ls, kp, side = build_workload(args.particles)
    log('workload: %d spheres, side %.0f nm, N_kappa %d' % (len(plist), side, len(kp)))
    inc = incidences()
    t0 = time.time()
    base = LinearSystem(plist, PlaneWave(WL, *inc[0]), ls, kp, solver_type='gmres', solver_tolerance=args.tol,
                        store_coupling_matrix=False, coupling_matrix_lookup_resolution=2.0, interpolator_kind='cubic',
                        precision='mixed')
    base.prepare()
    prep_s = time.time() - t0
    op = base.coupling_matrix
    t_diag = base.t_matrix.diagonal
    log('prepare %.2f s (table %.2f s), len_rho %d' % (prep_s, op.table_seconds, len(op.radial_distance_array)))

    # right-hand sides (setup, outside the timed region): step j solves incidences nb j .. nb j + nb - 1 of the
    # 64 distinct ones; T is the same for every incidence, rhs = T a
    nb = args.batch
    n_rhs = nb * (args.warmup + args.steps)
    if n_rhs > len(inc):
        raise SystemExit('at most %d distinct right-hand sides (warmup + steps <= %d)' % (len(inc), len(inc) // nb))
    rhs = np.stack([t_diag * PlaneWave(WL, *inc[j]).swe_coefficients(plist, ls) for j in range(n_rhs)])

    tune = autotune(op, ((True, 'half'), (True, 'double'), (False, 'single')), not args.no_autotune)
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
    half_terms, half_lin_ratio = op.half_interpolation()
    flops_16 = E * (4 * half_terms + 8 * 8)
    log('matvec FP32 %.2f ms, FP64 %.2f ms; 8-RHS FP16 %.2f ms, FP32 %.2f ms, FP64 %.2f ms'
        % (k32, k64, kb16, kb32, kb64))

    def solve(rows, precision, tol):
        x, st, ok = op.solve_master(t_diag, rhs[rows], tol, precision=precision, with_stats=True)
        return x, st, ok

    # the right-hand sides and T in HBM before the timed region (cuda.DeviceArray); the solutions stay there
    # until the region ends. Every library call returns after its stream's work is done, so the solves need no
    # further synchronization around the timed region.
    from smuthi_amd.cuda import DeviceArray
    rhs_dev = DeviceArray.from_host(rhs)
    t_dev = DeviceArray.from_host(t_diag)

    # timed region: batched solves at the requested inner precision, distinct incidences every step
    prec = args.precision
    for w in range(args.warmup):
        op.solve_master(t_dev, rhs_dev[nb * w:nb * w + nb], args.tol, precision=prec, with_stats=True)
    t0 = time.time()
    sols, stats = [], []
    for stp in range(args.warmup, args.warmup + args.steps):
        x, stt, _ = op.solve_master(t_dev, rhs_dev[nb * stp:nb * stp + nb], args.tol, precision=prec,
                                    with_stats=True)
        sols.append(x)
        stats.append(stt)
    elapsed = time.time() - t0
    timed = slice(nb * args.warmup, n_rhs)
    x_timed = np.concatenate([x.to_host() for x in sols])
    # the same steps from host arrays (PCIe copies of rhs, T and x inside each solve)
    t0 = time.time()
    for stp in range(args.warmup, args.warmup + args.steps):
        solve(slice(nb * stp, nb * stp + nb), prec, args.tol)
    host_elapsed = time.time() - t0
    # every timed right-hand side against an FP64 solve of it (FP64 GMRES to 1e-10)
    t0 = time.time()
    x_ref, ref_stats, _ = solve(timed, 'double', 1e-10)
    log('FP64 reference solves (tol 1e-10) of %d RHS: %.2f s' % (len(x_ref), time.time() - t0))
    errs = _errors(x_timed, x_ref)
    worst_res = max(s['relative_residual'] for s in stats)
    n_ok = sum(1 for e in errs if e <= 1e-5) if worst_res <= args.tol else 0
    value = n_ok / elapsed
    log('batched %s solves: %.2f s for %d RHS, worst residual %.2e, worst coefficient error vs FP64 %.2e'
        % (prec, elapsed, len(errs), worst_res, max(errs)))

    # baselines on the same incidences: plain FP64 GMRES, batched (8 RHS per launch) and one RHS at a time
    # (smuthi's way), and the FP32-inner refinement
    first = slice(nb * args.warmup, nb * args.warmup + nb)
    t0 = time.time()
    _, st64, _ = solve(first, 'double', args.tol)
    fp64_batched_s = time.time() - t0
    t0 = time.time()
    _, st64s, _ = op.solve_master(t_diag, rhs[nb * args.warmup], args.tol, precision='double', with_stats=True)
    fp64_single_s = time.time() - t0
    t0 = time.time()
    x32, st32, _ = solve(first, 'mixed', args.tol)
    mixed_s = time.time() - t0
    e32 = _errors(x32, x_ref[:nb])
    fp64_batched_rate = nb / fp64_batched_s
    log('FP64 GMRES batched %.2f s per %d RHS, single %.2f s; FP32-inner refinement %.2f s per %d RHS'
        % (fp64_batched_s, nb, fp64_single_s, mixed_s, nb))

    cpu = cfg1 = None
    if not args.no_cpu_baseline:
        t0 = time.time()
        cpu = cpu_baseline(op.lookup_table, op.radial_distance_array, len(plist), st64s['iterations'], workdir)
        log('cpu baseline %.1f s' % (time.time() - t0))
        t0 = time.time()
        cfg1 = cfg1_cpu_vs_gpu(workdir)
        log('cfg1 CPU vs GPU %.1f s' % (time.time() - t0))

    half = prec == 'mixed16'
    kern_ms = kb16 if half else (kb32 if prec == 'mixed' else kb64)
    flops_k = flops_16 if half else flops_8
    achieved = flops_k / (kern_ms * 1e-3) / 1e12
    peak = MFMA_FP16_PEAK_TFLOPS if half else (MFMA_FP32_PEAK_TFLOPS if prec == 'mixed' else MFMA_FP64_PEAK_TFLOPS)
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
                         'GMRES tol %.0e, %s-in-FP64 refinement, %d distinct plane waves per step (polar angles '
                         '0-60 deg x TE/TM)' % (len(plist), args.tol,
                                               'FP16' if half else ('FP32' if prec == 'mixed' else 'no'), nb)),
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
        'rhs_checked': len(errs), 'rhs_within_1e-5': n_ok,
        'inputs': 'right-hand sides and T resident in HBM (cuda.DeviceArray) when the timed region starts',
        'pcie_inclusive': {'rhs_solves_per_s': len(errs) / host_elapsed, 'ms_per_step': host_elapsed / args.steps * 1e3,
                           'note': 'the same steps from host numpy arrays'},
        'coef_err_vs_fp64_worst': max(errs), 'coef_err_vs_fp64_median': float(np.median(errs)),
        'worst_relative_residual': worst_res,
        'solve_stats': stats,
        'fp64_reference_stats': ref_stats,
        'fp64_gmres_batched': {'seconds_per_batch': fp64_batched_s, 'rhs_solves_per_s': fp64_batched_rate, **st64},
        'fp64_gmres_single_rhs': {'seconds': fp64_single_s, 'rhs_solves_per_s': 1.0 / fp64_single_s, **st64s},
        'mixed_fp32_inner': {'seconds_per_batch': mixed_s, 'rhs_solves_per_s': nb / mixed_s,
                             'coef_err_vs_fp64_worst': max(e32), **st32},
        'speedup_vs_fp64_gmres_batched': value / fp64_batched_rate,
        'speedup_vs_fp64_gmres_single_rhs': value * fp64_single_s,
        'prepare_seconds': prep_s,
        'table_seconds': op.table_seconds,
        'autotune': tune,
        'roofline': {'bound': ('instruction issue: VALU ~67 % and MFMA ~10 % of the SIMD cycles in the SQ counters '
                               '(they do not co-issue on gfx950)') if half else 'mfma',
                     'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s',
                     'frac': achieved / peak,
                     'traffic': measured_traffic('batch16' if half else 'batch'),
                     'kernel': ('batch_half_kernel<3,3,%d>' % half_terms if half else 'batch_slots_kernel<3,3,4>'),
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
        'cfg1_cpu_vs_gpu': cfg1,
        'device': dev, 'compute_units': n_cu, 'host': platform.node(), 'host_cpus': os.cpu_count(),
    }
    return json.dumps(out)


def main_precision(args):
    """Precision configuration at the headline geometry (10,000 spheres, cubic spline 2 nm): for l_max = m_max
    = 3 and 4, 8 distinct plane waves solved with FP32 coupling matvecs inside FP64 refinement (tol 1e-6) and
    with plain FP64 GMRES (tol 1e-10, the reference); reported: coefficient error of every right-hand side,
    extinction cross section from the FP32 pipeline (mixed solution, FP32 PWE assembly) against the FP64 one,
    and |E| of the scattered field on a 512 x 512 xz grid above the slab (PWE, 2000 kappa x 180 alpha) from
    the FP32 pipeline against the FP64 one. The FP64 reference is this GPU's FP64 path, itself pinned to the
    CPU oracle at small sizes (the oracle cannot run 300k unknowns)."""
    from smuthi_amd import cuda, ensembles, far_field as ffm, fields as flds
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.linear_system import LinearSystem
    from smuthi_amd.scattered_field import scattered_field_pwe
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    dev, n_cu = cuda.device_info()
    inc = incidences()
    nb = 8
    ls, kp = workload_medium()
    side = 60000.0 * np.sqrt(args.particles / 10000.0)
    n_grid, n_kappa, n_alpha = 512, 2000, 180
    X, Z = np.meshgrid(np.linspace(-side / 2, side / 2, n_grid), np.linspace(330.0, 3000.0, n_grid))
    kpw = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1.4e-3,
                                                  neff_imag=1e-2)[:n_kappa]
    alpha = np.linspace(0, 2 * np.pi, n_alpha)
    out = {'metric': 'errors of the FP32-matvec pipeline vs FP64 at %d spheres' % args.particles, 'variants': {},
           'data': 'synthetic', 'device': dev, 'compute_units': n_cu,
           'config': {'workload': ('%d spheres monolayer on glass (headline geometry), cubic spline 2 nm, %d plane '
                                   'waves, tol %.0e; near field on %dx%d xz points (z 330-3000 nm) from the scattered '
                                   'PWE with %d kappa x %d alpha' % (args.particles, nb, args.tol, n_grid, n_grid,
                                                                      len(kpw), n_alpha))}}
    for l_max in (3, 4):
        plist = ensembles.monolayer(args.particles, side, 200.0, seed=2, l_max=l_max)
        t0 = time.time()
        base = LinearSystem(plist, PlaneWave(WL, *inc[0]), ls, kp, solver_type='gmres', solver_tolerance=args.tol,
                            store_coupling_matrix=False, coupling_matrix_lookup_resolution=2.0,
                            interpolator_kind='cubic', precision='mixed')
        base.prepare()
        op = base.coupling_matrix
        t = base.t_matrix.diagonal
        log('l_max %d: %d unknowns, prepare %.1f s' % (l_max, op.shape[0], time.time() - t0))
        tune = autotune(op, ((True, 'single'), (True, 'double')), not args.no_autotune)
        waves = [PlaneWave(WL, *inc[j]) for j in range(0, 2 * nb, 2)]
        rhs = np.stack([t * w.swe_coefficients(plist, ls) for w in waves])
        t0 = time.time()
        x32, st32, ok32 = op.solve_master(t, rhs, args.tol, precision='mixed', with_stats=True)
        s32 = time.time() - t0
        # FP16-operand inner solves (the first call captures the solver's graphs: timed on the second)
        op.solve_master(t, rhs, args.tol, precision='mixed16')
        t0 = time.time()
        x16, st16, ok16 = op.solve_master(t, rhs, args.tol, precision='mixed16', with_stats=True)
        s16 = time.time() - t0
        t0 = time.time()
        x64, st64, ok64 = op.solve_master(t, rhs, 1e-10, precision='double', with_stats=True)
        s64 = time.time() - t0
        coef = _errors(x32, x64)
        coef16 = _errors(x16, x64)
        ext = []
        for r, w in enumerate(waves):
            e32 = ffm.extinction_cross_section(w, ffm._per_particle(plist, x32[r]), ls, precision='single')
            e64 = ffm.extinction_cross_section(w, ffm._per_particle(plist, x64[r]), ls, precision='double')
            ext.append(abs(e32 - e64) / abs(e64))
        # near field of right-hand side 0 in the top layer (air)
        fields = {}
        for name, x, prec in (('fp32', x32[0], 'single'), ('fp64', x64[0], 'double')):
            up, _ = scattered_field_pwe(WL, ffm._per_particle(plist, x), ls, 1, k_parallel=kpw,
                                        azimuthal_angles=alpha, precision=prec)
            ex, ey, ez = up.electric_field(X.ravel(), np.zeros(X.size), Z.ravel(), precision=prec)
            fields[name] = np.sqrt(np.abs(ex) ** 2 + np.abs(ey) ** 2 + np.abs(ez) ** 2)
        e_err = float(np.max(np.abs(fields['fp32'] - fields['fp64'])) / np.max(fields['fp64']))
        e_rms = float(np.linalg.norm(fields['fp32'] - fields['fp64']) / np.linalg.norm(fields['fp64']))
        log('l_max %d: mixed %.2f s (%d its), mixed16 %.2f s (%d its), FP64 %.2f s; coef err %.2e (mixed16 '
            '%.2e), ext err %.2e, |E| err %.2e (rms %.2e)'
            % (l_max, s32, st32['iterations'], s16, st16['iterations'], s64, max(coef), max(coef16), max(ext), e_err,
               e_rms))
        out['variants']['l_max_%d' % l_max] = {
            'unknowns': op.shape[0], 'mixed_seconds_per_8_rhs': s32, 'mixed_rhs_solves_per_s': nb / s32,
            'fp64_seconds_per_8_rhs_tol_1e-10': s64, 'mixed_stats': st32, 'fp64_stats': st64,
            'coef_err_vs_fp64_worst': max(coef), 'coef_err_vs_fp64_median': float(np.median(coef)),
            'extinction_rel_err_worst': max(ext), 'near_field_abs_E_max_err_rel_to_max': e_err,
            'near_field_abs_E_rms_rel_err': e_rms, 'autotune': tune,
            'mixed16_seconds_per_8_rhs': s16, 'mixed16_rhs_solves_per_s': nb / s16, 'mixed16_stats': st16,
            'mixed16_coef_err_vs_fp64_worst': max(coef16), 'mixed16_certified': bool(ok16),
            'matvec_ms_8rhs': dict(zip(('fp16', 'fp32', 'fp64'),
                                       (op.bench_matvec_batch(p_, iters=2)[0] for p_ in ('half', 'single', 'double')))),
            'batch_kernels': ('batch_half_kernel<%d,%d,%d> (%s), batch_slots_kernel<%d,%d,4> (%s), '
                              'batch_slots64_kernel<%d,%d,4> (%s)'
                              % (l_max, l_max, op.half_interpolation()[0], '1 part' if l_max == 3 else '3 tile parts',
                                 l_max, l_max, '1 K-part' if l_max == 3 else '3 K-parts', l_max, l_max,
                                 '2 K-parts' if l_max == 3 else '6 K-parts'))}
        del op, base
    out['value'] = out['variants']['l_max_3']['coef_err_vs_fp64_worst']
    out['unit'] = 'relative coefficient error'
    out['higher_is_better'] = False
    return json.dumps(out)


def main_batching(args):
    """Batching configuration: ensembles of uneven size (log-uniform) in a 1 um film (n = 1.8) on glass, heights
    U(150, 600) nm, one shared volume lookup table (Catmull-Rom, the reference's CUDA 'cubic'), 8 of the 64
    plane waves per step solved for every ensemble at once (FP32-inner refinement, FP64 certificate), every
    (ensemble, field) system of the first timed step checked against an FP64 solve."""
    from smuthi_amd import cuda, ensembles, fields as flds
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.layers import LayerSystem
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    dev, n_cu = cuda.device_info()
    sizes = ensembles.batching_sizes(args.ensembles, seed=4, low=args.min_size, high=args.max_size)
    t0 = time.time()
    ens = [ensembles.batching_ensemble(n, 400 + k) for k, n in enumerate(sizes)]
    ls = LayerSystem([0, 1000, 0], [1.52, 1.8, 1])
    kp = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1.8, 1], neff_resolution=1e-2)
    batch = EnsembleBatch(WL, ens, ls, k_parallel=kp, resolution=args.resolution, interpolator_kind='cubic',
                          cubic_interpolant='catmull-rom')
    prep = time.time() - t0
    log('%d ensembles (%d..%d spheres, %d total), table %.1f s, setup %.1f s'
        % (len(sizes), min(sizes), max(sizes), sum(sizes), batch.table_seconds, prep))
    inc = incidences()
    nb = 8
    n_fields = nb * (args.warmup + args.steps)
    t0 = time.time()
    rhs = batch.right_hand_sides([PlaneWave(WL, *inc[j]) for j in range(n_fields)])
    log('right-hand sides %.1f s' % (time.time() - t0))
    op = batch.operator
    inner = 'half' if args.precision == 'mixed16' else 'single'
    tune = autotune(op, ((True, inner), (True, 'double')), not args.no_autotune)
    k16 = op.bench_matvec_batch('half', iters=2)[0] if args.precision == 'mixed16' else None
    k32, tot32 = op.bench_matvec_batch('single', iters=2)
    k64, tot64 = op.bench_matvec_batch('double', iters=2)
    if args.matvecs_only:
        # the volume matvec kernels alone (no solves): a short run for kernel traces and counters
        log('8-RHS volume matvecs over all ensembles: FP16 %s ms, FP32 %.3f ms, FP64 %.3f ms'
            % ('%.3f' % k16 if k16 is not None else '-', k32, k64))
        return json.dumps({'metric': '8-RHS volume matvec time, all ensembles', 'value': k16 if k16 else k32,
                           'unit': 'ms', 'higher_is_better': False, 'n_gpus': 1,
                           'matvec_ms': {'fp16_8rhs_all_ensembles': k16, 'fp32_8rhs_all_ensembles': k32,
                                         'fp64_8rhs_all_ensembles': k64},
                           'particles': sum(sizes), 'ensembles': len(sizes), 'autotune': tune,
                           'device': dev, 'compute_units': n_cu})
    for w in range(args.warmup):
        batch.solve_with_stats(rhs[nb * w:nb * w + nb], tol=args.tol, precision=args.precision)
    t0 = time.time()
    xs, stats = [], []
    for stp in range(args.warmup, args.warmup + args.steps):
        x, st, ok = batch.solve_with_stats(rhs[nb * stp:nb * stp + nb], tol=args.tol, precision=args.precision)
        xs.append(x)
        stats.append(st)
    elapsed = time.time() - t0
    # the systems of the first timed step against FP64 GMRES to 1e-10 (every ensemble, 8 plane waves)
    checked = slice(nb * args.warmup, nb * args.warmup + nb)
    log('timed solves %.1f s; FP64 reference solves of the first timed step' % elapsed)
    t0 = time.time()
    x_ref, ref_st, _ = batch.solve_with_stats(rhs[checked], tol=1e-10, precision='double')
    ref_s = time.time() - t0
    x = xs[0]
    errs = [e for a, b in zip(batch.split(x), batch.split(x_ref)) for e in _errors(a, b)]
    worst_res = max(s['relative_residual'] for s in stats)
    # the rate counts every timed system when all checked ones are within 1e-5 and every residual certifies
    n_timed = nb * len(sizes) * args.steps
    n_ok = n_timed if (worst_res <= args.tol and max(errs) <= 1e-5) else 0
    t0 = time.time()
    _, st64, _ = batch.solve_with_stats(rhs[nb * args.warmup:nb * args.warmup + nb], tol=args.tol,
                                        precision='double')
    fp64_s = time.time() - t0
    systems_per_step = nb * len(sizes)
    value = n_ok / elapsed
    E = float(sum((30 * n) ** 2 for n in sizes))
    flops = E * (4 * 32 + 8 * nb)  # volume bicubic: 2 tables x 16 taps per complex entry
    half = args.precision == 'mixed16'
    k_inner = k16 if half else k32
    peak_inner = MFMA_FP16_PEAK_TFLOPS if half else MFMA_FP32_PEAK_TFLOPS
    achieved = flops / (k_inner * 1e-3) / 1e12
    log('batched solves: %.2f s for %d systems, worst residual %.2e, worst error %.2e; FP64 GMRES %.2f s per step'
        % (elapsed, len(errs), worst_res, max(errs), fp64_s))
    out = {
        'metric': 'RHS-solves/s (ensemble x plane wave systems), err<=1e-5 vs FP64',
        'value': value, 'unit': 'RHS-solves/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': elapsed / args.steps * 1e3, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': ('fp16 operands / fp32 accumulation coupling matvec (MFMA), fp32 sweep residuals, fp64 refinement/'
                  'Arnoldi/certifying residual' if args.precision == 'mixed16' else
                  'fp32 coupling matvec (MFMA), fp64 refinement/Arnoldi/certifying residual'), 'data': 'synthetic',
        'config': {'workload': ('%d ensembles of %d-%d spheres (log-uniform, seed 4) in a 1 um film (n=1.8) on glass, '
                                'z U(150,600) nm, one volume table (Catmull-Rom, %.0f nm), 8 plane waves per step '
                                'for every ensemble, tol %.0e' % (len(sizes), min(sizes), max(sizes),
                                                                   args.resolution, args.tol)),
                   'ensembles': len(sizes), 'sizes': sizes, 'particles': sum(sizes),
                   'systems_per_step': systems_per_step, 'resolution_nm': args.resolution,
                   'table_rho_rows': int(len(batch.rho_array)), 'table_z_rows': int(len(batch.sum_z_array)),
                   'n_kappa': int(len(kp))},
        'systems_timed': n_timed, 'systems_checked_vs_fp64': len(errs),
        'systems_checked_within_1e-5': sum(1 for e in errs if e <= 1e-5), 'coef_err_vs_fp64_worst': max(errs),
        'worst_relative_residual': worst_res, 'solve_stats': stats, 'fp64_reference_seconds': ref_s,
        'fp64_reference_stats': ref_st,
        'fp64_gmres_batched': {'seconds_per_step': fp64_s, 'rhs_solves_per_s': systems_per_step / fp64_s, **st64},
        'speedup_vs_fp64_gmres_batched': value * fp64_s / systems_per_step,
        'matvec_ms': {'fp16_8rhs_all_ensembles': k16, 'fp32_8rhs_all_ensembles': k32, 'fp32_8rhs_total': tot32,
                      'fp64_8rhs_all_ensembles': k64, 'fp64_8rhs_total': tot64},
        'roofline': {'bound': ('mfma by the algorithmic count; in practice the tap rows: about one pair per '
                               '(receiver chunk, rho, z) cell, so each pair reads its 16 tap rows per table'),
                     'achieved': achieved, 'peak': peak_inner, 'unit': 'TFLOP/s', 'frac': achieved / peak_inner,
                     'traffic': None,
                     'kernel': ('batch_half_kernel<3,3,4,true>' if half else 'batch_slots_kernel<3,3,4,true>') +
                               ' (8 launches: 2 tables x 4 z slices)',
                     'flops_per_matvec': flops},
        'table_seconds': batch.table_seconds, 'setup_seconds': prep, 'autotune': tune,
        'device': dev, 'compute_units': n_cu,
    }
    return json.dumps(out)


def main_learned(args):
    """Learned near-field preconditioner: polydisperse spheres (r 60-180 nm, n 1.4-2.4 + 0-0.05i) inside the
    500 nm film (n = 2.0) of air / film / glass, volume lookup (Catmull-Rom). The network (K = 16 neighbours,
    hidden 256, BF16 weights) is trained on the device on solver solutions of training ensembles (random and
    plane-wave right-hand sides), a validation ensemble picks the network or the untrained first-order
    structure, then a held-out ensemble is solved for 16 plane waves (8 polar angles x TE/TM) with plain
    GMRES, the first-order structure, the trained network and the validated choice (plain GMRES as
    fallback); every solution is certified by its FP64 residual and checked against an FP64 solve."""
    from smuthi_amd import cuda, ensembles, fields as flds
    from smuthi_amd.batch import EnsembleBatch
    from smuthi_amd.initial_field import PlaneWave
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.learned import LearnedPreconditioner, solve_with_fallback
    if not cuda.enable_gpu():
        raise SystemExit('no MI355X device')
    dev, n_cu = cuda.device_info()
    stack = ([0, 500, 0], [1.52, 2.0, 1])
    ls = LayerSystem(*stack)
    kp = flds.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=stack[1], neff_resolution=1e-2)
    rng = np.random.default_rng(5)

    def make(sizes, seed0):
        ens = [ensembles.learned_ensemble(n, seed0 + k) for k, n in enumerate(sizes)]
        return EnsembleBatch(WL, ens, ls, k_parallel=kp, resolution=args.resolution, interpolator_kind='cubic',
                             cubic_interpolant='catmull-rom')

    def prepare(lp, b):
        parts = b.particle_list
        return lp.prepare(b.operator, np.array([p.position for p in parts]), [p.radius for p in parts],
                          [p.refractive_index for p in parts], 0.0, 500.0, WL, b.t_diagonal,
                          ensemble_offsets=b.particle_offsets)

    lp = LearnedPreconditioner(K=16, hidden=256, l_max=3, seed=0, prior_weight=args.prior_weight)
    t0 = time.time()
    train = make([args.train_size] * args.train_ensembles, 500)
    _, feat = prepare(lp, train)
    th = np.deg2rad(np.linspace(0, 60, 8))
    fields = [PlaneWave(WL, np.pi - t, 0.2, pol) for t in th for pol in (0, 1)]
    n = train.n_unknowns
    R = np.concatenate([rng.standard_normal((8, n)) + 1j * rng.standard_normal((8, n)),
                        train.right_hand_sides(fields[::2])])
    log('training ensembles ready (%d spheres); solving %d right-hand sides' % (train.n_unknowns // 30, len(R)))
    X, stx, _ = train.solve_with_stats(R, tol=1e-8, precision='mixed')
    lp.add_training_ensemble(train.operator, feat, R, X)
    data_s = time.time() - t0
    t0 = time.time()
    L0 = lp.data_loss()[0]
    lp.train(args.train_steps, lr=1e-3)
    L1 = lp.data_loss()[0]
    train_s = time.time() - t0
    log('training data %.1f s, training %d steps %.1f s, loss %.4e -> %.4e' % (data_s, args.train_steps, train_s, L0,
                                                                             L1))
    del train
    # validation on a held-out ensemble: the network or the first-order structure
    val = make([args.val_size], 800)
    _, feat_v = prepare(lp, val)
    t0 = time.time()
    selection = lp.select(val.operator, feat_v, val.t_diagonal, val.right_hand_sides(fields[::4]), args.tol,
                          precision=args.precision)
    log('validation (%d spheres, 4 plane waves): %s in %.1f s' % (args.val_size, selection, time.time() - t0))
    del val
    test = make([args.test_size], 900)
    tune = autotune(test.operator, ((True, 'half' if args.precision == 'mixed16' else 'single'), (True, 'double')),
                    not args.no_autotune)
    _, feat_t = prepare(lp, test)
    rhs = test.right_hand_sides(fields)
    log('test ensemble (%d spheres) ready; FP64 reference solves of %d right-hand sides' % (args.test_size, len(rhs)))
    t0 = time.time()
    x_ref, _, _ = test.solve_with_stats(rhs, tol=1e-10, precision='double')
    log('FP64 reference %.1f s' % (time.time() - t0))
    res = {}
    for name in ('plain', 'first_order', 'learned', 'selected'):
        if name == 'first_order':
            lp.apply_prior(test.operator)
        elif name == 'learned':
            lp.apply(test.operator, feat_t)
        elif name == 'selected':
            lp.apply_selected(test.operator, feat_t)
        t0 = time.time()
        x, st, fb = solve_with_fallback(test.operator, test.t_diagonal, rhs, args.tol, precision=args.precision)
        secs = time.time() - t0
        errs = _errors(x, x_ref)
        res[name] = {'seconds': secs, 'rhs_solves_per_s': len(rhs) / secs, 'iterations': st['iterations'],
                     'fallback': bool(fb), 'worst_relative_residual': st['relative_residual'],
                     'coef_err_vs_fp64_worst': max(errs), 'stats': st}
        log('%s: %.2f s, %d iterations, fallback %s, error %.2e' % (name, secs, st['iterations'], fb, max(errs)))
    # plain FP64 GMRES on the GPU, no preconditioner: the 16 right-hand sides batched, and one right-hand
    # side per solve as smuthi runs them
    LearnedPreconditioner.remove(test.operator)
    t0 = time.time()
    _, st64b, _ = test.solve_with_stats(rhs, tol=args.tol, precision='double')
    s64b = time.time() - t0
    t0 = time.time()
    _, st64s, _ = test.solve_with_stats(rhs[:1], tol=args.tol, precision='double')
    s64s = time.time() - t0
    log('plain FP64 GMRES: %.2f s for %d right-hand sides batched, %.2f s for one' % (s64b, len(rhs), s64s))
    fp64 = {'batched': {'seconds': s64b, 'rhs_solves_per_s': len(rhs) / s64b, 'stats': st64b},
            'single_rhs': {'seconds': s64s, 'rhs_solves_per_s': 1.0 / s64s, 'stats': st64s}}
    value = res['selected']['rhs_solves_per_s'] if res['selected']['coef_err_vs_fp64_worst'] <= 1e-5 else 0.0
    out = {
        'metric': ('RHS-solves/s with the learned near-field preconditioner as selected on a validation ensemble '
                   '(held-out test ensemble), err<=1e-5 vs FP64'),
        'value': value, 'unit': 'RHS-solves/s', 'n_gpus': 1, 'higher_is_better': True, 'vs_baseline': None,
        'dtype': 'fp32 coupling matvec, fp64 refinement/Arnoldi/preconditioner/residual; network bf16 weights, '
                 'fp32 accumulation', 'data': 'synthetic',
        'config': {'workload': ('held-out ensemble of %d polydisperse spheres in a 500 nm film (n=2.0) of air/film/'
                                'glass, volume lookup (Catmull-Rom, %.0f nm), 16 plane waves, tol %.0e; trained on '
                                '%d ensembles of %d spheres x 16 right-hand sides'
                                % (args.test_size, args.resolution, args.tol, args.train_ensembles, args.train_size)),
                   'K': 16, 'hidden': 256, 'train_steps': args.train_steps},
        'iterations': {k: v['iterations'] for k, v in res.items()},
        'fallback_rate': sum(v['fallback'] for v in res.values() if v is not res['plain']) / 3.0,
        'speedup_vs_plain': res['selected']['rhs_solves_per_s'] / res['plain']['rhs_solves_per_s'],
        'plain_fp64_gmres': fp64,
        'speedup_vs_plain_fp64_gmres_batched': value / fp64['batched']['rhs_solves_per_s'],
        'speedup_vs_plain_fp64_gmres_single_rhs': value / fp64['single_rhs']['rhs_solves_per_s'],
        'validation': selection, 'prior_weight': args.prior_weight,
        'autotune': tune,
        'results': res, 'training': {'loss_untrained': L0, 'loss_trained': L1, 'seconds': train_s,
                                     'data_seconds': data_s, 'rows': args.train_size * args.train_ensembles,
                                     'solve_stats': stx},
        'device': dev, 'compute_units': n_cu,
    }
    return json.dumps(out)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=4)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--particles', type=int, default=10000)
    ap.add_argument('--tol', type=float, default=1e-6)
    ap.add_argument('--batch', type=int, default=8, help='right-hand sides solved together per step')
    ap.add_argument('--precision', default='mixed16', choices=['mixed16', 'mixed', 'double'],
                    help='inner matvec precision of the timed solves (mixed16: FP16 operands, mixed: FP32)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-autotune', action='store_true', help='keep the default launch shapes')
    ap.add_argument('--matvec-iters', type=int, default=5)
    ap.add_argument('--config', type=int, default=2, choices=[2, 3, 4, 5],
                    help='2: the 10k-sphere solve (default); 4: ensembles batched; 5: learned preconditioner')
    ap.add_argument('--ensembles', type=int, default=32)
    ap.add_argument('--min-size', type=int, default=100)
    ap.add_argument('--max-size', type=int, default=1000)
    ap.add_argument('--resolution', type=float, default=5.0)
    ap.add_argument('--train-ensembles', type=int, default=4)
    ap.add_argument('--train-size', type=int, default=800)
    ap.add_argument('--test-size', type=int, default=2500)
    ap.add_argument('--train-steps', type=int, default=400)
    ap.add_argument('--val-size', type=int, default=800)
    ap.add_argument('--prior-weight', type=float, default=1e-3,
                    help='pull of the learned scales towards the first-order structure')
    ap.add_argument('--matvecs-only', action='store_true',
                    help='config 4: time the volume matvec kernels and stop before the solves')
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
    if args.config == 3:
        return main_precision(args)
    if args.config == 4:
        return main_batching(args)
    if args.config == 5:
        return main_learned(args)
    return main_headline(args)


if __name__ == '__main__':
    # the drop-in LinearSystem reports 'Number of unknowns' on stdout like smuthi does;
    # keep stdout to the one JSON line
    _stdout = sys.stdout
    with contextlib.redirect_stdout(sys.stderr):
        _line = main()
    if _line is not None:
        print(_line, file=_stdout, flush=True)
