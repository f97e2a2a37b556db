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


class DeviceArray:
    """A complex128 array in the GPU's HBM (smh_device_alloc), C-contiguous, 1-D or 2-D.

    For callers that keep right-hand sides and solutions on the device:
    ``solve_master`` of the GPU operators takes DeviceArrays and returns one.
    Row slices ``a[i:j]`` and single rows ``a[i]`` of a 2-D array are views
    (they keep the owner alive)."""

    itemsize = 16

    def __init__(self, shape, _ptr=None, _owner=None):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        if len(self.shape) not in (1, 2) or min(self.shape) < 1:
            raise ValueError('DeviceArray shapes are [n] or [rows, n] with positive sizes')
        self._owner = _owner
        if _ptr is None:
            p = ctypes.c_void_p()
            _lib.check(_lib.load().smh_device_alloc(context(), self.nbytes, ctypes.byref(p)), context(),
                       'smh_device_alloc')
            self.ptr = p.value
            self._owns = True
        else:
            self.ptr = _ptr
            self._owns = False

    @property
    def ndim(self):
        return len(self.shape)

    @property
    def size(self):
        n = 1
        for s in self.shape:
            n *= s
        return n

    @property
    def nbytes(self):
        return self.size * self.itemsize

    @classmethod
    def from_host(cls, a):
        import numpy as np
        h = np.ascontiguousarray(a, dtype=np.complex128)
        d = cls(h.shape)
        _lib.check(_lib.load().smh_device_copy(context(), d.ptr, h.ctypes.data, d.nbytes), context(),
                   'smh_device_copy')
        return d

    @classmethod
    def empty(cls, shape):
        return cls(shape)

    def to_host(self):
        import numpy as np
        h = np.empty(self.shape, dtype=np.complex128)
        _lib.check(_lib.load().smh_device_copy(context(), h.ctypes.data, self.ptr, self.nbytes), context(),
                   'smh_device_copy')
        return h

    def __getitem__(self, key):
        if self.ndim != 2:
            raise IndexError('only rows of 2-D DeviceArrays can be taken')
        rows, n = self.shape
        owner = self._owner or self
        if isinstance(key, slice):
            start, stop, step = key.indices(rows)
            if step != 1 or stop <= start:
                raise IndexError('row slices must be contiguous and non-empty')
            return DeviceArray((stop - start, n), _ptr=self.ptr + start * n * self.itemsize, _owner=owner)
        i = int(key)
        if i < 0:
            i += rows
        if not 0 <= i < rows:
            raise IndexError('row %d out of range' % key)
        return DeviceArray((n,), _ptr=self.ptr + i * n * self.itemsize, _owner=owner)

    def __del__(self):
        if getattr(self, '_owns', False) and self.ptr and _ctx is not None and _lib._LIB is not None:
            _lib.load().smh_device_free(_ctx, ctypes.c_void_p(self.ptr))
            self.ptr = None
