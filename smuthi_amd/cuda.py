"""GPU switch with smuthi's interface (utility/cuda.py:5-68).

``use_gpu`` selects the GPU operators in LinearSystem and the PWE field
evaluation; ``enable_gpu()`` probes the MI355X through libsmh and, like smuthi,
warns and leaves ``use_gpu = False`` when no usable device is found. The
device context is created once per process and shared by all operators.
"""
import atexit
import ctypes
import warnings

from . import _lib

use_gpu = False
default_blocksize = 128  # accepted for API compatibility; the kernels pick their own launch shape

_ctx = None
_device = 0


def enable_gpu(enable=True, device=0):
    """Enable the GPU path if a gfx950 device can be initialised."""
    global use_gpu, _device
    if not enable:
        use_gpu = False
        return False
    try:
        _device = device
        context()
        use_gpu = True
    except Exception as e:  # mirrors smuthi's warn-and-fall-back behaviour
        warnings.warn('Unable to enable the MI355X path: %s' % e)
        use_gpu = False
    return use_gpu


def context():
    """The process-wide libsmh context (created on first use)."""
    global _ctx
    if _ctx is None:
        lib = _lib.load()
        handle = ctypes.c_void_p()
        _lib.check(lib.smh_init(_device, ctypes.byref(handle)), None, 'smh_init(device=%d)' % _device)
        _ctx = handle
        atexit.register(_release)
    return _ctx


def device_info():
    lib = _lib.load()
    name = ctypes.create_string_buffer(64)
    cu = ctypes.c_int()
    _lib.check(lib.smh_device_info(context(), name, 64, ctypes.byref(cu)), context(), 'smh_device_info')
    return name.value.decode(), cu.value


def _release():
    global _ctx
    if _ctx is not None:
        _lib.load().smh_destroy(_ctx)
        _ctx = None
