"""taco_tensor_t descriptors over device buffers.

A descriptor is host memory (the struct and its small arrays); pos, crd and
vals are device pointers.  ``Desc`` keeps every buffer it points to alive for
as long as the descriptor itself.  Device memory comes from torch (plumbing
only: allocation and the stream), the kernels come from libtaco_hip.
"""
import ctypes
from ctypes import POINTER, byref, c_int, c_int32, c_uint8

from . import _lib

DENSE, SPARSE = 0, 1  # taco_mode_t

CSIZE = {"f64": 64, "f32": 32, "bf16": 16}


class Desc:
    """Owns a ``taco_tensor_t`` and references to the buffers it points at."""

    def __init__(self, dims, mode_types, csize, keep=()):
        order = len(dims)
        self._dims = (c_int32 * max(order, 1))(*dims)
        self._ordering = (c_int32 * max(order, 1))(*range(order))
        self._types = (c_int * max(order, 1))(*mode_types)
        self._level_arrays = []
        self._indices = (POINTER(POINTER(c_uint8)) * max(order, 1))()
        for k, mt in enumerate(mode_types):
            arr = (POINTER(c_uint8) * (2 if mt == SPARSE else 1))()
            self._level_arrays.append(arr)
            self._indices[k] = ctypes.cast(arr, POINTER(POINTER(c_uint8)))
        self._fill = (c_uint8 * 8)()
        self.t = _lib.TacoTensorT()
        self.t.order = order
        self.t.dimensions = self._dims
        self.t.csize = csize
        self.t.mode_ordering = self._ordering
        self.t.mode_types = self._types
        self.t.indices = self._indices
        self.t.fill_value = self._fill
        self.t.vals_size = 0
        self.keep = list(keep)

    @property
    def ptr(self):
        return byref(self.t)

    def set_level(self, level, arrays):
        for q, a in enumerate(arrays):
            self._level_arrays[level][q] = ctypes.cast(ctypes.c_void_p(a), POINTER(c_uint8))

    def set_vals(self, addr, count):
        self.t.vals = ctypes.cast(ctypes.c_void_p(addr), POINTER(c_uint8))
        self.t.vals_size = int(count)

    def level_ptr(self, level, q):
        p = self._level_arrays[level][q]
        return ctypes.cast(p, ctypes.c_void_p).value or 0

    def vals_ptr(self):
        return ctypes.cast(self.t.vals, ctypes.c_void_p).value or 0


def _addr(t):
    return t.data_ptr() if t is not None and t.numel() > 0 else 0


def csr(pos, crd, vals, shape, dtype="f64"):
    """CSR descriptor over device tensors pos (int32, m+1), crd (int32),
    vals (float64/float32, or int16/uint16 bit patterns for bf16)."""
    m, n = shape
    d = Desc([m, n], [DENSE, SPARSE], CSIZE[dtype], keep=(pos, crd, vals))
    d._level_arrays[0][0] = ctypes.cast(ctypes.pointer(c_int32(m)), POINTER(c_uint8))
    d.keep.append(d._level_arrays[0][0])
    d.set_level(1, [_addr(pos), _addr(crd)])
    d.set_vals(_addr(vals), crd.numel())
    return d


def csr_empty(shape, dtype="f64"):
    """CSR result descriptor whose arrays an assemble function allocates."""
    m, n = shape
    d = Desc([m, n], [DENSE, SPARSE], CSIZE[dtype])
    d._msize = c_int32(m)
    d._level_arrays[0][0] = ctypes.cast(ctypes.pointer(d._msize), POINTER(c_uint8))
    return d


def dense(t, dtype="f64"):
    """Dense row-major descriptor over a device tensor (1-D or 2-D)."""
    shape = list(t.shape)
    d = Desc(shape, [DENSE] * len(shape), CSIZE[dtype], keep=(t,))
    d._sizes = [c_int32(s) for s in shape]
    for k, s in enumerate(d._sizes):
        d._level_arrays[k][0] = ctypes.cast(ctypes.pointer(s), POINTER(c_uint8))
    d.set_vals(_addr(t), t.numel())
    return d
