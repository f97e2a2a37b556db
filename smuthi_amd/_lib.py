"""ctypes binding of libsmh.so (declared in include/smh.h).

The library is built in-tree by ``smuthi_amd.build``. Loading fails loudly
when it is missing: there is no CPU fallback for the GPU operators.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libsmh.so')

SMH_OK = 0
SMH_ERR_ARG = 1
SMH_ERR_HIP = 2
SMH_ERR_UNSUPPORTED = 3
SMH_ERR_STATE = 4
SMH_ERR_NOCONV = 5

INTERP_LINEAR = 0
INTERP_CATMULL_ROM = 1
INTERP_SPLINE = 2

FP32 = 0
FP64 = 1
FP16 = 2

SOLVE_FP64 = 0
SOLVE_MIXED = 1
SOLVE_MIXED16 = 2
SOLVE_LU = 3

_ERR_NAMES = {1: 'invalid argument', 2: 'HIP error', 3: 'unsupported', 4: 'invalid state', 5: 'not converged'}

# name -> (restype, argtypes); every symbol declared in include/smh.h
_c = ctypes
_p = _c.c_void_p
_i64 = _c.c_int64
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)
_fp = _c.POINTER(_c.c_float)
SIGNATURES = {
    'smh_init': (_c.c_int, [_c.c_int, _c.POINTER(_p)]),
    'smh_destroy': (_c.c_int, [_p]),
    'smh_last_error': (_c.c_char_p, [_p]),
    'smh_device_info': (_c.c_int, [_p, _c.c_char_p, _c.c_int, _ip]),
    'smh_radial_create': (_c.c_int, [_p, _i64, _dp, _ip, _ip, _c.c_int, _c.c_int, _c.c_double, _c.c_double, _i64,
                                     _c.c_int, _c.POINTER(_p)]),
    'smh_volume_create': (_c.c_int, [_p, _i64, _dp, _ip, _ip, _c.c_int, _c.c_int, _c.c_double, _c.c_double,
                                     _c.c_double, _c.c_double, _i64, _i64, _c.c_int, _c.POINTER(_p)]),
    'smh_radial_create_ensembles': (_c.c_int, [_p, _c.c_int, _c.POINTER(_i64), _dp, _ip, _ip, _c.c_int, _c.c_int,
                                               _c.c_double, _c.c_double, _i64, _c.c_int, _c.POINTER(_p)]),
    'smh_volume_create_ensembles': (_c.c_int, [_p, _c.c_int, _c.POINTER(_i64), _dp, _ip, _ip, _c.c_int, _c.c_int,
                                               _c.c_double, _c.c_double, _c.c_double, _c.c_double, _i64, _i64,
                                               _c.c_int, _c.POINTER(_p)]),
    'smh_op_set_preconditioner': (_c.c_int, [_p, _c.c_int, _ip, _dp]),
    'smh_op_near_blocks': (_c.c_int, [_p, _c.c_int, _ip, _dp]),
    'smh_op_download_near_blocks': (_c.c_int, [_p, _dp]),
    'smh_op_set_near_scaling': (_c.c_int, [_p, _fp]),
    'smh_mlp_create': (_c.c_int, [_p, _c.c_int, _c.c_int, _c.c_int, _c.c_uint64, _c.POINTER(_p)]),
    'smh_mlp_destroy': (_c.c_int, [_p]),
    'smh_mlp_num_params': (_i64, [_p]),
    'smh_mlp_get_params': (_c.c_int, [_p, _fp]),
    'smh_mlp_set_params': (_c.c_int, [_p, _fp]),
    'smh_mlp_forward': (_c.c_int, [_p, _i64, _fp, _fp]),
    'smh_mlp_backward': (_c.c_int, [_p, _i64, _fp, _fp, _fp]),
    'smh_mlp_add_data': (_c.c_int, [_p, _p, _fp, _i64, _dp, _dp]),
    'smh_mlp_data_loss': (_c.c_int, [_p, _fp, _fp, _dp]),
    'smh_mlp_set_prior_weight': (_c.c_int, [_p, _c.c_double]),
    'smh_mlp_train': (_c.c_int, [_p, _c.c_int, _c.c_double, _dp]),
    'smh_volume_build_table': (_c.c_int, [_p, _c.c_double, _c.c_double, _c.c_double, _c.c_double, _dp, _i64, _dp,
                                          _dp, _dp]),
    'smh_op_destroy': (_c.c_int, [_p]),
    'smh_op_upload_table': (_c.c_int, [_p, _dp, _dp]),
    'smh_radial_build_table': (_c.c_int, [_p, _c.c_double, _c.c_double, _c.c_double, _dp, _i64, _dp, _dp]),
    'smh_op_download_table': (_c.c_int, [_p, _dp, _dp]),
    'smh_op_matvec': (_c.c_int, [_p, _dp, _dp, _i64, _c.c_int]),
    'smh_device_alloc': (_c.c_int, [_p, _i64, _c.POINTER(_c.c_void_p)]),
    'smh_device_free': (_c.c_int, [_p, _c.c_void_p]),
    'smh_device_copy': (_c.c_int, [_p, _c.c_void_p, _c.c_void_p, _i64]),
    'smh_op_matvec_noalloc': (_c.c_int, [_p, _c.c_void_p, _c.c_void_p, _c.c_int]),
    'smh_op_bench_matvec': (_c.c_int, [_p, _c.c_int, _c.c_int, _dp, _dp]),
    'smh_op_bench_matvec_batch': (_c.c_int, [_p, _c.c_int, _c.c_int, _dp, _dp]),
    'smh_op_half_interp': (_c.c_int, [_p, _ip, _dp]),
    'smh_op_autotune': (_c.c_int, [_p, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _ip, _ip, _dp, _ip]),
    'smh_solve': (_c.c_int, [_p, _dp, _dp, _dp, _i64, _c.c_double, _c.c_int, _c.c_int, _c.c_int, _dp]),
    'smh_pwe_field': (_c.c_int, [_p, _c.c_int, _i64, _dp, _dp, _i64, _dp, _dp, _c.c_double, _c.c_double, _i64, _dp,
                                 _dp, _c.c_int, _dp]),
    'smh_pwe_assemble': (_c.c_int, [_p, _i64, _dp, _c.c_int, _c.c_int, _dp, _c.c_double, _i64, _dp, _dp, _dp, _i64,
                                    _dp, _dp, _c.c_int, _dp]),
    'smh_swe_field': (_c.c_int, [_p, _i64, _dp, _dp, _dp, _c.c_int, _c.c_int, _dp, _c.c_double, _c.c_double, _i64,
                                 _dp, _c.c_int, _c.c_double, _c.c_int, _dp, _dp]),
    'smh_explicit_create': (_c.c_int, [_p, _i64, _dp, _ip, _ip, _c.c_int, _c.c_int, _ip, _c.c_int, _dp, _dp,
                                       _c.c_int, _dp, _i64, _dp, _dp, _dp, _dp, _c.POINTER(_p)]),
    'smh_explicit_blocks': (_c.c_int, [_p, _i64, _i64, _dp, _c.c_int, _c.c_int, _ip, _c.c_int, _dp, _dp, _c.c_int,
                                       _dp, _i64, _dp, _dp, _dp, _dp, _dp, _dp]),
    'smh_pwe_to_swe': (_c.c_int, [_p, _i64, _dp, _dp, _dp, _i64, _dp, _dp, _dp, _dp, _c.c_int, _c.c_int, _i64, _dp,
                                  _dp, _dp]),
    'smh_bench_copy': (_c.c_int, [_p, _i64, _dp]),
    'smh_host_bessel_j': (_c.c_int, [_c.c_int, _i64, _dp, _dp]),
    'smh_host_spherical_hankel': (_c.c_int, [_c.c_int, _i64, _dp, _dp]),
}


class SmhError(RuntimeError):
    """Raised for a nonzero return code of a libsmh call."""

    def __init__(self, code, message):
        self.code = code
        super().__init__('%s (%s)' % (message, _ERR_NAMES.get(code, 'error %d' % code)))


_LIB = None


def load():
    """Load libsmh.so once; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError('libsmh.so is not built; run `python -m smuthi_amd.build` '
                           '(hipcc, gfx950) before using the GPU operators')
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


DPTR = _dp  # pointer type of the float64 / complex128 arguments (also for device addresses)


def dptr(a):
    """Pointer to the data of a C-contiguous float64 / complex128 array."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS'] and a.dtype in (np.float64, np.complex128), (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def i64ptr(a):
    assert a.flags['C_CONTIGUOUS'] and a.dtype == np.int64
    return a.ctypes.data_as(_c.POINTER(_i64))


def fptr(a):
    """Pointer to the data of a C-contiguous float32 array."""
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS'] and a.dtype == np.float32, (a.dtype, a.flags)
    return a.ctypes.data_as(_fp)


def iptr(a):
    if a is None:
        return None
    assert a.flags['C_CONTIGUOUS'] and a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def check(code, ctx=None, what='libsmh call'):
    if code != SMH_OK:
        msg = what
        if ctx is not None:
            err = load().smh_last_error(ctx)
            if err:
                msg = '%s: %s' % (what, err.decode(errors='replace'))
        raise SmhError(code, msg)
