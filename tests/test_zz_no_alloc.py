"""smh_op_matvec_noalloc (the matvec that allocates nothing: device vectors in,
staging in the library's fixed scratch array) against smh_op_matvec, the
existing entry point, on two calls in a row, the second with a smaller
operator; FP32 identical, FP64 to rounding of the LDS accumulation order."""
import ctypes

import numpy as np
import pytest

from oracle import basics as ob

pytestmark = pytest.mark.gpu

WL = 550.0


@pytest.fixture(scope='module')
def gpu():
    from smuthi_amd import cuda
    assert cuda.enable_gpu(), 'MI355X not available'
    return cuda


def _hip():
    hip = ctypes.CDLL('libamdhip64.so')
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


def _noalloc(op, x, precision):
    from smuthi_amd import _lib
    from smuthi_amd import cuda
    hip = _hip()
    nbytes = x.size * 16
    dx, dy = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dx), nbytes) == 0 and hip.hipMalloc(ctypes.byref(dy), nbytes) == 0
    try:
        assert hip.hipMemcpy(dx, x.ctypes.data, nbytes, 1) == 0
        code = _lib.load().smh_op_matvec_noalloc(op._op, dx, dy, precision)
        _lib.check(code, cuda.context(), 'smh_op_matvec_noalloc')
        out = np.empty_like(x)
        assert hip.hipMemcpy(out.ctypes.data, dy, nbytes, 2) == 0
        return out
    finally:
        hip.hipFree(dx)
        hip.hipFree(dy)


def test_noalloc_matvec_matches_matvec(gpu):
    from smuthi_amd import _lib
    from smuthi_amd import ensembles
    from smuthi_amd.layers import LayerSystem
    from smuthi_amd.linear_system import CouplingMatrixRadialLookupCUDA
    kp = ob.reasonable_Sommerfeld_kpar_contour(WL, layer_refractive_indices=[1.52, 1], neff_resolution=1e-2)
    ls = LayerSystem([0, 0], [1.52, 1])
    big = ensembles.monolayer(400, 12000.0, 200.0, seed=21)
    small = ensembles.monolayer(30, 3000.0, 200.0, seed=22, l_max=3, m_max=2)
    rng = np.random.default_rng(7)
    for plist in (big, small):
        print('no-alloc matvec: %d spheres' % len(plist), flush=True)
        op = CouplingMatrixRadialLookupCUDA(WL, plist, ls, k_parallel=kp, resolution=5.0, interpolator_kind='cubic')
        x = rng.standard_normal(op.shape[0]) + 1j * rng.standard_normal(op.shape[0])
        x = np.ascontiguousarray(x)
        want32 = op.matvec(x, 'single')
        got32 = _noalloc(op, x, _lib.FP32)
        assert np.array_equal(got32, want32), np.abs(got32 - want32).max()
        want64 = op.matvec(x, 'double')
        got64 = _noalloc(op, x, _lib.FP64)
        assert np.abs(got64 - want64).max() <= 1e-14 * np.abs(want64).max()
