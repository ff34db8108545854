"""Helpers of the GPU tests: host CSR -> device buffers -> descriptors."""
import numpy as np
import torch

from taco_amd import _lib, descriptors as D


def dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to("cuda")


def bf16_bits(a):
    """float64 array -> bf16 bit patterns (round to nearest even) as int16."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.view(torch.int16)


def bf16_values(a):
    """The float64 values of a after rounding to bf16."""
    t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(torch.bfloat16)
    return t.to(torch.float64).numpy()


def csr_desc(A, dtype="f64"):
    pos = dev(A.pos)
    crd = dev(A.crd)
    if dtype == "f64":
        vals = dev(A.vals, torch.float64)
    elif dtype == "f32":
        vals = dev(A.vals, torch.float32)
    else:
        vals = bf16_bits(A.vals).to("cuda")
    return D.csr(pos, crd, vals, A.shape, dtype)


def spmv(A, x, sched=None, dtype="f64"):
    tdt = torch.float64 if dtype == "f64" else torch.float32
    a = csr_desc(A, dtype)
    xd = dev(x, tdt)
    y = torch.full((A.m,), float("nan"), dtype=tdt, device="cuda")
    rc = _lib.lib().taco_hip_spmv_compute(D.dense(y, dtype).ptr, a.ptr,
                                          D.dense(xd, dtype).ptr, sched)
    _lib.check(rc, "taco_hip_spmv_compute")
    _lib.check(_lib.lib().taco_hip_synchronize())
    return y.cpu().numpy().astype(np.float64)


def spmm(A, B, sched=None, dtype="f64"):
    if dtype == "bf16":
        a = csr_desc(A, "bf16")
        bd = bf16_bits(B).to("cuda")
        bdesc = D.dense(bd, "bf16")
        C = torch.full((A.m, B.shape[1]), float("nan"), dtype=torch.float32, device="cuda")
        cdesc = D.dense(C, "f32")
    else:
        tdt = torch.float64 if dtype == "f64" else torch.float32
        a = csr_desc(A, dtype)
        bd = dev(B, tdt)
        bdesc = D.dense(bd, dtype)
        C = torch.full((A.m, B.shape[1]), float("nan"), dtype=tdt, device="cuda")
        cdesc = D.dense(C, dtype)
    rc = _lib.lib().taco_hip_spmm_compute(cdesc.ptr, a.ptr, bdesc.ptr, sched)
    _lib.check(rc, "taco_hip_spmm_compute")
    _lib.check(_lib.lib().taco_hip_synchronize())
    return C.cpu().numpy().astype(np.float64)


def spgemm(A, B):
    import ctypes
    a = csr_desc(A)
    b = csr_desc(B)
    c = D.csr_empty((A.m, B.n))
    L = _lib.lib()
    _lib.check(L.taco_hip_spgemm_assemble(c.ptr, a.ptr, b.ptr), "assemble")
    _lib.check(L.taco_hip_spgemm_compute(c.ptr, a.ptr, b.ptr), "compute")
    _lib.check(L.taco_hip_synchronize())
    nnz = c.t.vals_size
    pos = np.empty(A.m + 1, np.int32)
    crd = np.empty(nnz, np.int32)
    vals = np.empty(nnz, np.float64)
    for host, dptr in ((pos, c.level_ptr(1, 0)), (crd, c.level_ptr(1, 1)),
                       (vals, c.vals_ptr())):
        if host.nbytes:
            _lib.check(L.taco_hip_memcpy(host.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(dptr), host.nbytes, 1))
    for dptr in (c.level_ptr(1, 0), c.level_ptr(1, 1), c.vals_ptr()):
        L.taco_hip_unified_free(ctypes.c_void_p(dptr))
    return pos, crd, vals


def sched(algo, block=0, items=0, group=0, rows=0):
    s = _lib.Schedule()
    s.algo, s.block_size, s.items_per_thread = algo, block, items
    s.group_size, s.rows_per_group = group, rows
    return s
