"""Device blocks freed through taco_hip_unified_free are kept for reuse when
they are 1 MiB or larger (sizes rounded up to 2 MiB): the next allocation of
the same rounded size gets the same block back, without hipMalloc.  Smaller
blocks go straight back to the runtime.  Each block still holds what is
written to it."""
import ctypes

import numpy as np
import pytest

from taco_amd import _lib

pytestmark = pytest.mark.gpu


def _write_read(p, n, seed):
    L = _lib.lib()
    h = (np.arange(n, dtype=np.int64) * 2654435761 + seed).astype(np.int32)
    _lib.check(L.taco_hip_memcpy(ctypes.c_void_p(p), h.ctypes.data_as(ctypes.c_void_p),
                                 h.nbytes, 0))
    g = np.zeros_like(h)
    _lib.check(L.taco_hip_memcpy(g.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p),
                                 g.nbytes, 1))
    assert np.array_equal(g, h)


def test_large_block_reused_after_free():
    L = _lib.lib()
    assert L.taco_hip_should_use_unified_memory() == 0
    n = 3 * (1 << 20) // 4 + 17                  # 3 MiB + 68 B: the 4 MiB class
    p = L.taco_hip_unified_alloc(4 * n)
    assert p
    _write_read(p, n, 1)
    L.taco_hip_unified_free(ctypes.c_void_p(p))
    q = L.taco_hip_unified_alloc(4 * n - 100)    # same class: the same block
    assert q == p
    _write_read(q, n, 2)
    r = L.taco_hip_unified_alloc(4 * n)          # the block is in use: a new one
    assert r and r != q
    _write_read(r, n, 3)
    s = L.taco_hip_unified_alloc(9 << 20)        # another class
    assert s and s not in (q, r)
    for b in (q, r, s):
        L.taco_hip_unified_free(ctypes.c_void_p(b))


def test_blocks_freed_between_spgemm_assembles_are_reused():
    """Two assembles in a row, the first's C freed in between: the second's
    C arrays come from the cache, and its pattern is the first's."""
    import torch
    from taco_amd import descriptors as D, workloads as W
    A, B = W.csr_fixed(60000, 50000, 8, 81), W.csr_fixed(50000, 40000, 8, 82)
    ad = D.csr(torch.from_numpy(A.pos).cuda(), torch.from_numpy(A.crd).cuda(),
               torch.from_numpy(A.vals).cuda(), A.shape)
    bd = D.csr(torch.from_numpy(B.pos).cuda(), torch.from_numpy(B.crd).cuda(),
               torch.from_numpy(B.vals).cuda(), B.shape)
    L = _lib.lib()
    outs = []
    for _ in range(2):
        c = D.csr_empty((A.m, B.n))
        _lib.check(L.taco_hip_spgemm_assemble(c.ptr, ad.ptr, bd.ptr))
        nnz = c.t.vals_size
        assert nnz * 4 >= (1 << 20)
        crd = np.empty(nnz, np.int32)
        _lib.check(L.taco_hip_memcpy(crd.ctypes.data_as(ctypes.c_void_p),
                                     ctypes.c_void_p(c.level_ptr(1, 1)), crd.nbytes, 1))
        ptrs = (c.level_ptr(1, 0), c.level_ptr(1, 1), c.vals_ptr())
        outs.append((ptrs, crd))
        for p in ptrs:
            L.taco_hip_unified_free(ctypes.c_void_p(p))
    assert outs[1][0][1:] == outs[0][0][1:]   # crd and vals blocks reused
    assert np.array_equal(outs[0][1], outs[1][1])
