"""Batched uneven SpMV / SpMM workloads and the persistent batched launch.

A ``BatchedWorkload`` holds many CSR instances concatenated in device memory
(64-bit instance offsets, int32 indices inside an instance), generated from a
seed: per instance, rows log-uniform in [rmin, rmax] (square), per-row entry
counts uniform in [kmin, kmax], and one of three structures — uniform random
columns, banded (columns within +-64 of the diagonal) or block diagonal with
dense 8x8 blocks (every row holds its whole block).  Values, x and B are
f32 U[-1,1] generated on the device.

``Batch`` wraps the persistent kernel: per-instance schedules (lanes per
row), launches over the whole batch or a range of instances, per-instance
event timing for the autotuner, and hipGraph capture.
"""
import ctypes

import numpy as np

from . import _lib
from . import workloads as W

STREAM_INST_ROWS = 21
STREAM_INST_STRUCT = 22
STREAM_BATCH_VALS = 23
STREAM_BATCH_X = 24
STREAM_BATCH_B = 25

STRUCT_NAMES = {W.STRUCT_UNIFORM: "uniform", W.STRUCT_BANDED: "banded",
                W.STRUCT_BLOCK: "block8"}


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def instance_plan(ninst, seed, rmin, rmax):
    """Rows and structure of each instance (deterministic in seed)."""
    u = W.uniform(ninst, seed, STREAM_INST_ROWS, 0.0, 1.0)
    m = np.floor(rmin * (rmax / rmin) ** u).astype(np.int64)
    m = np.clip(m, rmin, rmax).astype(np.int32)
    s = np.floor(W.uniform(ninst, seed, STREAM_INST_STRUCT, 0.0, 3.0)).astype(np.int32)
    return m, np.clip(s, 0, 2)


def instance_pattern(i, m, structure, seed, kmin, kmax):
    """pos/crd of instance i (host arrays)."""
    iseed = (seed * 1_000_003 + i) & 0xFFFFFFFFFFFF
    if structure == W.STRUCT_BLOCK:
        rows = np.arange(m, dtype=np.int64)
        counts = np.minimum(8, m - (rows // 8) * 8).astype(np.int32)
        band = 8
    else:
        counts = W.uniform_counts(m, kmin, kmax, iseed)
        band = 64
    A = W.csr_from_counts(m, m, counts, iseed, structure, band, values=False)
    return A.pos, A.crd


class BatchedWorkload:
    def __init__(self, ninst=4096, rmin=1000, rmax=200000, kmin=2, kmax=64,
                 seed=4, width=32, keep_host=False, spmm=True):
        import torch
        self.ninst, self.seed, self.width = ninst, seed, width
        self.m, self.structure = instance_plan(ninst, seed, rmin, rmax)
        self.kmin, self.kmax = kmin, kmax
        patterns = [instance_pattern(i, int(self.m[i]), int(self.structure[i]),
                                     seed, kmin, kmax) for i in range(ninst)] if keep_host else None
        # First pass: entry counts, to size the device buffers.
        if patterns is None:
            self.nnz = np.zeros(ninst, np.int64)
        else:
            self.nnz = np.array([p[1].size for p in patterns], np.int64)
        self.host = patterns
        self.pos_off = np.concatenate([[0], np.cumsum(self.m.astype(np.int64) + 1)])
        self.row_off = np.concatenate([[0], np.cumsum(self.m.astype(np.int64))])
        total_rows = int(self.row_off[-1])
        dev = "cuda"
        self.pos = torch.empty(int(self.pos_off[-1]), dtype=torch.int32, device=dev)
        if patterns is None:
            # Generate and upload instance by instance; sizes become known as
            # we go, so the crd buffer grows in chunks and is packed at the end.
            crd_parts = []
            for i in range(ninst):
                pos, crd = instance_pattern(i, int(self.m[i]), int(self.structure[i]),
                                            seed, kmin, kmax)
                self.nnz[i] = crd.size
                self.pos[self.pos_off[i]:self.pos_off[i + 1]].copy_(torch.from_numpy(pos))
                crd_parts.append(torch.from_numpy(crd).to(dev))
            self.crd = torch.cat(crd_parts) if crd_parts else torch.empty(0, dtype=torch.int32, device=dev)
            del crd_parts
        else:
            for i, (pos, _) in enumerate(patterns):
                self.pos[self.pos_off[i]:self.pos_off[i + 1]].copy_(torch.from_numpy(pos))
            self.crd = torch.from_numpy(np.concatenate([p[1] for p in patterns])).to(dev)
        self.ent_off = np.concatenate([[0], np.cumsum(self.nnz)])
        total = int(self.ent_off[-1])
        L = _lib.lib()
        self.vals = torch.empty(max(total, 1), dtype=torch.float32, device=dev)
        self.x = torch.empty(max(total_rows, 1), dtype=torch.float32, device=dev)
        self.y = torch.zeros(max(total_rows, 1), dtype=torch.float32, device=dev)
        _lib.check(L.taco_hip_gen_uniform_f32_device(total, seed, STREAM_BATCH_VALS, -1.0, 1.0,
                                                     ctypes.c_void_p(self.vals.data_ptr())))
        _lib.check(L.taco_hip_gen_uniform_f32_device(total_rows, seed, STREAM_BATCH_X, -1.0, 1.0,
                                                     ctypes.c_void_p(self.x.data_ptr())))
        self.B = self.C = None
        if spmm:
            self.B = torch.empty(max(total_rows * width, 1), dtype=torch.float32, device=dev)
            self.C = torch.zeros(max(total_rows * width, 1), dtype=torch.float32, device=dev)
            _lib.check(L.taco_hip_gen_uniform_f32_device(total_rows * width, seed, STREAM_BATCH_B,
                                                         -1.0, 1.0, ctypes.c_void_p(self.B.data_ptr())))
        torch.cuda.synchronize()
        self.total_nnz = total
        self.total_rows = total_rows

    def instance(self, i):
        """Host CSR of instance i with its f32 values widened to f64, and its
        x and B segments (for checking)."""
        a, b = int(self.ent_off[i]), int(self.ent_off[i + 1])
        r0, r1 = int(self.row_off[i]), int(self.row_off[i + 1])
        pos = self.pos[self.pos_off[i]:self.pos_off[i + 1]].cpu().numpy()
        crd = self.crd[a:b].cpu().numpy()
        vals = self.vals[a:b].cpu().numpy().astype(np.float64)
        A = W.Csr(r1 - r0, r1 - r0, pos, crd, vals)
        x = self.x[r0:r1].cpu().numpy().astype(np.float64)
        B = None
        if self.B is not None:
            B = self.B[r0 * self.width:r1 * self.width].cpu().numpy().astype(np.float64).reshape(-1, self.width)
        return A, x, B

    def y_of(self, i):
        r0, r1 = int(self.row_off[i]), int(self.row_off[i + 1])
        return self.y[r0:r1].cpu().numpy().astype(np.float64)

    def c_of(self, i):
        r0, r1 = int(self.row_off[i]), int(self.row_off[i + 1])
        return self.C[r0 * self.width:r1 * self.width].cpu().numpy().astype(np.float64).reshape(-1, self.width)

    def algorithmic_bytes(self, op):
        """Compulsory traffic: pos, crd, vals once; x/y (SpMV) or B/C (SpMM)."""
        base = 4 * int(self.pos_off[-1]) + 8 * self.total_nnz
        if op == 0:
            return base + 8 * self.total_rows
        return base + 8 * self.total_rows * self.width


class Batch:
    def __init__(self, wl, op, sched, rows_per_task=256):
        L = _lib.lib()
        self.wl, self.op = wl, op
        self._p = ctypes.c_void_p()
        sched = np.ascontiguousarray(sched, dtype=np.int32)
        m = np.ascontiguousarray(wl.m, dtype=np.int32)
        nnz = np.ascontiguousarray(wl.nnz, dtype=np.int64)
        inp = wl.x if op == 0 else wl.B
        out = wl.y if op == 0 else wl.C
        _lib.check(L.taco_hip_batch_create(
            op, wl.ninst, _p(m), _p(m), _p(nnz), _p(sched), rows_per_task, wl.width,
            ctypes.c_void_p(wl.pos.data_ptr()), ctypes.c_void_p(wl.crd.data_ptr()),
            ctypes.c_void_p(wl.vals.data_ptr()), ctypes.c_void_p(inp.data_ptr()),
            ctypes.c_void_p(out.data_ptr()), ctypes.byref(self._p)), "batch create")

    def set_schedules(self, sched):
        sched = np.ascontiguousarray(sched, dtype=np.int32)
        _lib.check(_lib.lib().taco_hip_batch_set_schedules(self._p, _p(sched)))

    def run(self):
        _lib.check(_lib.lib().taco_hip_batch_run(self._p), "batch run")

    def run_range(self, i0, i1):
        _lib.check(_lib.lib().taco_hip_batch_run_range(self._p, i0, i1), "batch range run")

    def time_instances(self, reps=3):
        """Per-instance time (ns) of one launch over that instance alone."""
        t = np.zeros(self.wl.ninst)
        _lib.check(_lib.lib().taco_hip_batch_time_instances(self._p, reps, _p(t)),
                   "batch instance timing")
        return t

    def capture(self):
        _lib.check(_lib.lib().taco_hip_batch_capture(self._p), "batch capture")

    def launch_graph(self):
        _lib.check(_lib.lib().taco_hip_batch_launch_graph(self._p), "batch graph launch")

    def info(self):
        out = (ctypes.c_int64 * 3)()
        _lib.check(_lib.lib().taco_hip_batch_info(self._p, out))
        return {"instances": out[0], "tasks": out[1], "workgroups": out[2]}

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None and _lib._lib is not None:
            _lib._lib.taco_hip_batch_destroy(self._p)
            self._p = None


# Per-instance SpMV schedules: lanes per row, or 0 for the stream schedule.
SPMV_LANES = (1, 2, 4, 8, 16, 32, 64, 0)
SPMM_LANES = (8, 16, 32, 64)


def default_schedule(op, ninst):
    """The untuned choice: thread per row for SpMV (taco's default GPU
    schedule), one 32-lane group per row for SpMM."""
    return np.full(ninst, 1 if op == 0 else 32, np.int32)
