"""CPU oracle of the sparse hot path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker (or the timed CPU baseline).  The
product (taco_amd) never imports it.

Two libraries:
  liboracle.so      oracle/taco_cpu.c + oracle/oracle_exact.c: a C/OpenMP
                    restatement of taco's generated CPU code, and the
                    extended-precision (long double + Neumaier) "FP64
                    oracle" with condition sums for the tolerance gate.
  _ref/libtaco_ref.so  the real taco library compiled from the reference's own
                    sources (oracle/Makefile, target `ref`), driven by
                    oracle/ref_driver.cpp.  Present only where it was built.

Parity status of each function is pinned by tests/test_oracle.py against the
reference's known-answer tests, the golden fixtures in tests/golden (written
by the real taco), closed forms and brute force.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libtaco_ref.so")

_c = None
_ref = None

i32p = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def lib():
    global _c
    if _c is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _c = ctypes.CDLL(LIB)
        _c.oracle_spgemm_assemble.restype = ctypes.c_int64
    return _c


def ref_available():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_LIB)
        _ref.taco_ref_last_error.restype = ctypes.c_char_p
        _ref.taco_ref_last_source.restype = ctypes.c_char_p
    return _ref


def set_threads(n):
    lib().oracle_set_num_threads(int(n))


# ---- restatement of taco's generated CPU code ----

def spmv(A, x, dtype=np.float64):
    y = np.empty(A.m, dtype=dtype)
    f = lib().oracle_spmv_f64 if dtype == np.float64 else lib().oracle_spmv_f32
    vals = np.ascontiguousarray(A.vals, dtype=dtype)
    xx = np.ascontiguousarray(x, dtype=dtype)
    f(A.m, A.n, _p(A.pos), _p(A.crd), _p(vals), _p(xx), _p(y))
    return y


def spmm(A, B, dtype=np.float64):
    k = B.shape[1]
    C = np.empty((A.m, k), dtype=dtype)
    f = lib().oracle_spmm_f64 if dtype == np.float64 else lib().oracle_spmm_f32
    vals = np.ascontiguousarray(A.vals, dtype=dtype)
    BB = np.ascontiguousarray(B, dtype=dtype)
    f(A.m, A.n, k, _p(A.pos), _p(A.crd), _p(vals), _p(BB), _p(C))
    return C


def spgemm_assemble(A, B):
    """Returns (pos, crd) of C = A B (taco's two-phase assemble)."""
    pos = np.empty(A.m + 1, dtype=np.int32)
    nnz = lib().oracle_spgemm_assemble(A.m, B.n, _p(A.pos), _p(A.crd),
                                       _p(B.pos), _p(B.crd), _p(pos), None)
    if nnz < 0:
        raise MemoryError("oracle_spgemm_assemble failed")
    crd = np.empty(int(nnz), dtype=np.int32)
    lib().oracle_spgemm_assemble(A.m, B.n, _p(A.pos), _p(A.crd), _p(B.pos),
                                 _p(B.crd), _p(pos), _p(crd))
    return pos, crd


def spgemm_compute(A, B, pos, crd):
    vals = np.empty(len(crd), dtype=np.float64)
    lib().oracle_spgemm_compute_f64(A.m, B.n, _p(A.pos), _p(A.crd), _p(A.vals),
                                    _p(B.pos), _p(B.crd), _p(B.vals), _p(pos),
                                    _p(crd), _p(vals))
    return vals


# ---- extended-precision oracle ----

def exact_spmv(A, x):
    y = np.empty(A.m)
    cond = np.empty(A.m)
    lib().exact_spmv(A.m, _p(A.pos), _p(A.crd), _p(np.ascontiguousarray(A.vals, dtype=np.float64)),
                     _p(np.ascontiguousarray(x, dtype=np.float64)), _p(y), _p(cond))
    return y, cond


def exact_spmm(A, B):
    k = B.shape[1]
    C = np.empty((A.m, k))
    cond = np.empty((A.m, k))
    lib().exact_spmm(A.m, k, _p(A.pos), _p(A.crd),
                     _p(np.ascontiguousarray(A.vals, dtype=np.float64)),
                     _p(np.ascontiguousarray(B, dtype=np.float64)), _p(C), _p(cond))
    return C, cond


def exact_spgemm_values(A, B, pos, crd):
    vals = np.empty(len(crd))
    cond = np.empty(len(crd))
    lib().exact_spgemm_values(A.m, B.n, _p(A.pos), _p(A.crd), _p(A.vals),
                              _p(B.pos), _p(B.crd), _p(B.vals), _p(pos),
                              _p(crd), _p(vals), _p(cond))
    return vals, cond


# ---- error metrics against the FP64 oracle ----

def error_report(got, exact, cond, tol):
    """Condition-scaled gate |got - exact| <= tol * cond (cond = sum |a*b|),
    plus max elementwise-relative error and the normwise ||d||inf/||y||inf."""
    got = np.asarray(got, dtype=np.float64).ravel()
    exact = np.asarray(exact, dtype=np.float64).ravel()
    cond = np.asarray(cond, dtype=np.float64).ravel()
    diff = np.abs(got - exact)
    # Entries with cond == 0 must be exact zeros.
    scaled = np.where(cond > 0, diff / np.where(cond > 0, cond, 1.0),
                      np.where(diff > 0, np.inf, 0.0))
    nz = np.abs(exact) > 0
    rel = np.zeros_like(diff)
    rel[nz] = diff[nz] / np.abs(exact[nz])
    ymax = np.max(np.abs(exact)) if exact.size else 0.0
    worst = int(np.argmax(scaled)) if scaled.size else -1
    return {
        "max_scaled": float(scaled.max()) if scaled.size else 0.0,
        "worst_index": worst,
        "max_rel": float(rel.max()) if rel.size else 0.0,
        "normwise": float(diff.max() / ymax) if ymax > 0 else float(diff.max() if diff.size else 0.0),
        "tol": tol,
        "ok": bool((scaled <= tol).all()) and bool(np.isfinite(got).all()),
    }


# ---- the real taco (oracle/_ref) ----

def ref_spmv(A, x, threads=1, reps=0):
    r = ref()
    r.taco_ref_set_num_threads(int(threads))
    y = np.empty(A.m)
    t = np.zeros(max(reps, 1))
    rc = r.taco_ref_spmv_f64(A.m, A.n, _p(A.pos), _p(A.crd),
                             _p(np.ascontiguousarray(A.vals, dtype=np.float64)),
                             _p(np.ascontiguousarray(x, dtype=np.float64)),
                             _p(y), int(reps), _p(t))
    if rc:
        raise RuntimeError(r.taco_ref_last_error().decode())
    return (y, t[:reps]) if reps else y


def ref_spmm(A, B, threads=1, reps=0):
    r = ref()
    r.taco_ref_set_num_threads(int(threads))
    k = B.shape[1]
    C = np.empty((A.m, k))
    t = np.zeros(max(reps, 1))
    rc = r.taco_ref_spmm_f64(A.m, A.n, k, _p(A.pos), _p(A.crd),
                             _p(np.ascontiguousarray(A.vals, dtype=np.float64)),
                             _p(np.ascontiguousarray(B, dtype=np.float64)),
                             _p(C), int(reps), _p(t))
    if rc:
        raise RuntimeError(r.taco_ref_last_error().decode())
    return (C, t[:reps]) if reps else C


def ref_spgemm(A, B, threads=1, reps=0):
    r = ref()
    r.taco_ref_set_num_threads(int(threads))
    op, oc, ov = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    ta = np.zeros(max(reps, 1))
    tc = np.zeros(max(reps, 1))
    rc = r.taco_ref_spgemm_f64(A.m, A.n, B.n, _p(A.pos), _p(A.crd), _p(A.vals),
                               _p(B.pos), _p(B.crd), _p(B.vals),
                               ctypes.byref(op), ctypes.byref(oc),
                               ctypes.byref(ov), int(reps), _p(ta), _p(tc))
    if rc:
        raise RuntimeError(r.taco_ref_last_error().decode())
    pos = np.ctypeslib.as_array(ctypes.cast(op, ctypes.POINTER(ctypes.c_int32)),
                                (A.m + 1,)).copy()
    nnz = int(pos[-1])
    crd = np.ctypeslib.as_array(ctypes.cast(oc, ctypes.POINTER(ctypes.c_int32)),
                                (max(nnz, 1),))[:nnz].copy()
    vals = np.ctypeslib.as_array(ctypes.cast(ov, ctypes.POINTER(ctypes.c_double)),
                                 (max(nnz, 1),))[:nnz].copy()
    for p in (op, oc, ov):
        r.taco_ref_free(p)
    if reps:
        return pos, crd, vals, ta[:reps], tc[:reps]
    return pos, crd, vals
