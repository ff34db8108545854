"""Tensor files: taco's read() and write() (reference include/taco/
storage/file_io_mtx.h, file_io_tns.h; src/storage/file_io_mtx.cpp,
file_io_tns.cpp) for MatrixMarket (.mtx, .ttx) and FROSTT (.tns).

The files are parsed by the library (taco_hip_read_tensor_file) into a
coordinate list.  ``read`` packs it into a facade Tensor of the requested
format (CSR or dense, order 1 or 2): every coordinate is inserted and the
tensor packed, so repeated coordinates are summed, as taco's pack does.
``read_coordinates`` returns the list itself for any order.  ``write``
writes a facade Tensor (or a coordinate list) back.
"""
import ctypes

import numpy as np

from . import _lib


class Coordinates:
    """A tensor file's contents: dims, 0-based coords [nnz, order], values,
    and whether the file was a dense MatrixMarket array."""

    def __init__(self, dims, coords, vals, dense=False):
        self.dims = [int(d) for d in dims]
        self.coords = np.ascontiguousarray(coords, np.int32).reshape(-1, len(self.dims))
        self.vals = np.ascontiguousarray(vals, np.float64)
        self.dense = bool(dense)

    @property
    def order(self):
        return len(self.dims)

    def packed(self):
        """(coords, vals) sorted by coordinate (first mode slowest) with
        repeated coordinates summed in file order: the components taco's
        pack stores."""
        if len(self.vals) == 0:
            return self.coords, self.vals
        keys = np.ravel_multi_index(tuple(self.coords.T.astype(np.int64)), self.dims)
        order = np.argsort(keys, kind="stable")
        k = keys[order]
        first = np.concatenate([[True], k[1:] != k[:-1]])
        starts = np.flatnonzero(first)
        vals = np.add.reduceat(self.vals[order], starts) if len(starts) else self.vals[:0]
        return self.coords[order][starts], vals


def read_coordinates(path):
    L = _lib.lib()
    h = ctypes.c_void_p()
    _lib.check(L.taco_hip_read_tensor_file(str(path).encode(), ctypes.byref(h)), "read")
    try:
        order, nnz, dense = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int32()
        dims = (ctypes.c_int32 * 16)()
        _lib.check(L.taco_hip_coo_info(h, ctypes.byref(order), dims, 16, ctypes.byref(nnz),
                                       ctypes.byref(dense)))
        if order.value > 16:
            raise _lib.TacoHipError(f"{path}: order {order.value} above 16")
        coords = np.empty((nnz.value, order.value), np.int32)
        vals = np.empty(nnz.value, np.float64)
        _lib.check(L.taco_hip_coo_data(h, coords.ctypes.data_as(ctypes.c_void_p),
                                       vals.ctypes.data_as(ctypes.c_void_p)))
        return Coordinates(list(dims)[:order.value], coords, vals, dense.value)
    finally:
        L.taco_hip_coo_destroy(h)


def read(path, fmt=None, name=None, dtype=None):
    """taco::read(filename, format): a packed facade Tensor (order 1 or 2).
    fmt defaults to CSR for matrices and dense for vectors."""
    from . import api
    c = read_coordinates(path)
    if c.order not in (1, 2):
        raise api.TacoError(f"{path}: order {c.order}; the facade holds vectors and matrices "
                            "(read_coordinates returns any order)")
    if fmt is None:
        fmt = api.CSR if c.order == 2 else api.dense_vector
    t = api.Tensor(name or "A", c.dims, fmt, dtype or api.Float64)
    coords, vals = c.packed()
    if fmt == api.CSR:
        m = c.dims[0]
        counts = np.bincount(coords[:, 0], minlength=m) if len(vals) else np.zeros(m, np.int64)
        pos = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        t.set_csr(pos, coords[:, 1], vals)
    else:
        d = np.zeros(c.dims, np.float64)
        if len(vals):
            d[tuple(coords.T)] = vals
        t.set_values(d)
    return t


def write(path, tensor):
    """taco::write(filename, tensor) for a facade Tensor or Coordinates:
    MatrixMarket coordinate (sparse) or array (dense) by the tensor's format,
    FROSTT coordinates for .tns."""
    from . import api
    L = _lib.lib()
    if isinstance(tensor, Coordinates):
        dims, coords, vals, dense = tensor.dims, tensor.coords, tensor.vals, tensor.dense
    elif tensor.format == api.CSR:
        pos, crd, vals = tensor.to_csr()
        rows = np.repeat(np.arange(len(pos) - 1, dtype=np.int32), np.diff(pos))
        dims, coords, vals = tensor.dims, np.stack([rows, crd], 1), vals
        dense = False
    else:
        a = np.asarray(tensor.to_numpy(), np.float64)
        dims = tensor.dims
        if str(path).endswith(".tns"):
            idx = np.argwhere(np.ones(a.shape, bool))
            coords, vals, dense = idx, a.reshape(-1), False
        else:
            # MatrixMarket arrays list values with the first mode fastest
            coords, vals, dense = np.zeros((0, len(dims)), np.int32), a.reshape(-1, order="F"), True
    coords = np.ascontiguousarray(coords, np.int32)
    vals = np.ascontiguousarray(vals, np.float64)
    d = (ctypes.c_int32 * len(dims))(*dims)
    _lib.check(L.taco_hip_write_tensor_file(str(path).encode(), len(dims), d, len(vals),
                                            coords.ctypes.data_as(ctypes.c_void_p),
                                            vals.ctypes.data_as(ctypes.c_void_p),
                                            1 if dense else 0), "write")
