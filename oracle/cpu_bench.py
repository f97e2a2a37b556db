"""CPU baseline runner for bench.py (test infrastructure: run only as the
benchmark's cpu_baseline leg, in a child process that never touches the GPU).

Times the oracle's restatement of smuthi's CPU radial-lookup matvec
(CouplingMatrixRadialLookupCPU.matvec, linear_system.py:903-975: per (n1, n2)
a scipy interp1d of the table column, gather, phase, dense dot) on all host
cores -- the 30 row groups n1 split over a fork pool, one BLAS thread per
worker -- and, for the 50-sphere configuration, the reference's full CPU
solve: scipy GMRES (restart 20, x0 = rhs, tol) on 1 - T W with that matvec
(linear_system.py:271-286).

    python -m oracle.cpu_bench matvec TABLE.npy RHO.npy POS.npy KIND WORKERS
    python -m oracle.cpu_bench solve TABLE.npy RHO.npy PARTICLES.npz KIND TOL
prints one JSON line.
"""
import json
import multiprocessing
import os
import sys
import time

import numpy as np
import scipy.sparse.linalg

from oracle import coupling as oc
from oracle.system import Sphere

_STATE = {}


def _blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max([d.get('num_threads', 1) for d in threadpool_info() if d.get('user_api') == 'blas'] or [1])
    except Exception:  # threadpoolctl missing or no BLAS found
        return None


def _worker(rows):
    from threadpoolctl import threadpool_limits
    with threadpool_limits(1):
        table = np.load(_STATE['table'], mmap_mode='r')
        rho = np.load(_STATE['rho'])
        pos = np.load(_STATE['pos'])
        plist = [Sphere(list(p), 1.5 + 0.01j, 100.0, 3, 3) for p in pos]
        loop = oc.RadialLookupCPUReferenceLoop(np.asarray(table), rho, plist, _STATE['kind'], n1_rows=rows)
        v = np.random.default_rng(1).standard_normal(30 * len(pos)) + 0j
        t0 = time.time()
        loop.matvec(v, n1_subset=rows)
        return time.time() - t0


def matvec(table, rho, pos, kind, workers):
    _STATE.update(table=table, rho=rho, pos=pos, kind=kind)
    groups = [list(range(k, 30, workers)) for k in range(workers)]
    groups = [g for g in groups if g]
    ctx = multiprocessing.get_context('fork')
    with ctx.Pool(len(groups)) as pool:
        times = pool.map(_worker, groups)
    return dict(matvec_seconds=max(times), workers=len(groups), blas_threads_per_worker=1,
                row_group_seconds=[float(t) for t in times], cpu_count=os.cpu_count(),
                blas_threads_default=_blas_threads())


def solve(table, rho, particles, kind, tol):
    d = np.load(particles)
    plist = [Sphere(list(p), complex(n), float(r), 3, 3) for p, n, r in zip(d['pos'], d['n'], d['r'])]
    loop = oc.RadialLookupCPUReferenceLoop(np.load(table), np.load(rho), plist, kind)
    t = d['t']
    rhs = d['rhs']
    n = len(rhs)
    op = scipy.sparse.linalg.LinearOperator((n, n), dtype=complex, matvec=lambda v: v - t * loop.matvec(v))
    its = [0]

    def cb(_):
        its[0] += 1

    t0 = time.time()
    x, info = scipy.sparse.linalg.gmres(op, rhs, x0=rhs, rtol=tol, atol=0.0, restart=20, callback=cb,
                                        callback_type='pr_norm')
    secs = time.time() - t0
    res = float(np.linalg.norm(rhs - op.matvec(x)) / np.linalg.norm(rhs))
    np.save(particles.replace('.npz', '_x.npy'), x)
    return dict(seconds=secs, iterations=its[0], info=int(info), relative_residual=res, cores=1,
                cpu_count=os.cpu_count(), blas_threads_default=_blas_threads())


if __name__ == '__main__':
    mode = sys.argv[1]
    if mode == 'matvec':
        out = matvec(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], int(sys.argv[6]))
    else:
        out = solve(sys.argv[2], sys.argv[3], sys.argv[4], sys.argv[5], float(sys.argv[6]))
    print(json.dumps(out), flush=True)
