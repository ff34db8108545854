"""ctypes binding of libtaco_hip.so (the C ABI declared in include/taco_hip.h).

The library is built in-tree by ``make -C taco_amd/csrc`` (which the
repository's ``__graft_entry__.build()`` runs).  There is no
fallback: if the shared object is missing or does not load, every entry
point raises, so a run can never silently compute on the CPU.
"""
import ctypes
import os
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int,
                    c_int32, c_int64, c_size_t, c_uint8, c_uint64, c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtaco_hip.so")


class TacoTensorT(Structure):
    """Mirror of ``taco_tensor_t`` (include/taco/taco_tensor_t.h)."""
    _fields_ = [
        ("order", c_int32),
        ("dimensions", POINTER(c_int32)),
        ("csize", c_int32),
        ("mode_ordering", POINTER(c_int32)),
        ("mode_types", POINTER(c_int)),
        ("indices", POINTER(POINTER(POINTER(c_uint8)))),
        ("vals", POINTER(c_uint8)),
        ("fill_value", POINTER(c_uint8)),
        ("vals_size", c_int32),
    ]


class Schedule(Structure):
    """Mirror of ``taco_hip_schedule``."""
    _fields_ = [
        ("algo", c_int32),
        ("block_size", c_int32),
        ("items_per_thread", c_int32),
        ("group_size", c_int32),
        ("rows_per_group", c_int32),
        ("reserved", c_int32 * 3),
    ]


# taco_hip_algo
THREAD_PER_ROW, GROUP_PER_ROW, NNZ_SPLIT, MERGE_PATH, MFMA = range(5)
ALGO_NAMES = {THREAD_PER_ROW: "thread_per_row", GROUP_PER_ROW: "group_per_row",
              NNZ_SPLIT: "nnz_split", MERGE_PATH: "merge_path", MFMA: "mfma"}

P_T = POINTER(TacoTensorT)

# name -> (restype, argtypes)
_SIGS = {
    "taco_hip_should_use_codegen": (c_int, []),
    "taco_hip_set_codegen_enabled": (None, [c_int]),
    "taco_hip_should_use_unified_memory": (c_int, []),
    "taco_hip_set_unified_memory_enabled": (None, [c_int]),
    "taco_hip_get_default_compiler_flags": (c_char_p, []),
    "taco_hip_unified_alloc": (c_void_p, [c_size_t]),
    "taco_hip_unified_free": (None, [c_void_p]),
    "taco_hip_unified_realloc": (c_void_p, [c_void_p, c_size_t, c_size_t]),
    "taco_hip_unified_clear": (c_int, [c_void_p, c_size_t]),
    "taco_hip_last_error": (c_char_p, []),
    "taco_hip_device_count": (c_int, []),
    "taco_hip_device_arch": (c_char_p, []),
    "taco_hip_set_stream": (c_int, [c_void_p]),
    "taco_hip_get_stream": (c_void_p, []),
    "taco_hip_set_blocking": (None, [c_int]),
    "taco_hip_synchronize": (c_int, []),
    "taco_hip_memcpy": (c_int, [c_void_p, c_void_p, c_size_t, c_int]),
    "taco_hip_memset": (c_int, [c_void_p, c_int, c_size_t]),
    "taco_hip_gen_uniform_f64": (None, [c_int64, c_uint64, c_uint64, c_double,
                                        c_double, c_void_p]),
    "taco_hip_gen_row_counts_zipf": (c_int64, [c_int32, c_double, c_int32,
                                               c_int32, c_double, c_uint64,
                                               c_void_p]),
    "taco_hip_gen_row_counts_uniform": (c_int64, [c_int32, c_int32, c_int32,
                                                  c_uint64, c_void_p]),
    "taco_hip_counts_to_pos": (c_int, [c_int32, c_void_p, c_void_p]),
    "taco_hip_gen_csr_columns": (c_int, [c_int32, c_int32, c_void_p, c_int32,
                                         c_int32, c_uint64, c_void_p]),
    "taco_hip_default_schedule": (None, [c_int32, c_int32, c_int64, c_int32,
                                         POINTER(Schedule)]),
    "taco_hip_spmv_compute": (c_int, [P_T, P_T, P_T, POINTER(Schedule)]),
    "taco_hip_spmv_compute_noalloc": (c_int, [P_T, P_T, P_T, POINTER(Schedule)]),
    "taco_hip_spmm_compute": (c_int, [P_T, P_T, P_T, POINTER(Schedule)]),
    "taco_hip_spgemm_assemble": (c_int, [P_T, P_T, P_T]),
    "taco_hip_spgemm_compute": (c_int, [P_T, P_T, P_T]),
    "init_taco_tensor_t": (P_T, [c_int32, c_int32, POINTER(c_int32),
                                 POINTER(c_int32), POINTER(c_int), c_void_p]),
    "deinit_taco_tensor_t": (None, [P_T]),
    "taco_hip_spmv_plan_create": (c_int, [P_T, c_int32, c_int32, POINTER(c_void_p)]),
    "taco_hip_spmv_plan_compute": (c_int, [c_void_p, P_T, P_T, c_int32]),
    "taco_hip_spmv_plan_destroy": (None, [c_void_p]),
    "taco_hip_spmv_plan_info": (c_int, [c_void_p, POINTER(c_int64)]),
    "taco_hip_spmv_plan_time_kernel": (c_int, [c_void_p, c_int32]),
    "taco_hip_spmv_plan_configure": (c_int, [c_void_p, c_int32, c_int32]),
    "taco_hip_spmv_plan_shape": (c_int, [c_void_p, POINTER(c_int32), POINTER(c_int32),
                                         POINTER(c_int32)]),
    "taco_hip_spmv_plan_kernel_time": (c_int, [c_void_p, POINTER(c_double), POINTER(c_int64)]),
    "taco_hip_batch_create": (c_int, [c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "taco_hip_batch_set_schedules": (c_int, [c_void_p, c_void_p]),
    "taco_hip_batch_run": (c_int, [c_void_p]),
    "taco_hip_batch_run_range": (c_int, [c_void_p, c_int32, c_int32]),
    "taco_hip_batch_time_instances": (c_int, [c_void_p, c_int32, c_void_p]),
    "taco_hip_batch_capture": (c_int, [c_void_p]),
    "taco_hip_batch_launch_graph": (c_int, [c_void_p]),
    "taco_hip_batch_destroy": (None, [c_void_p]),
    "taco_hip_batch_info": (c_int, [c_void_p, POINTER(c_int64)]),
    "taco_hip_gen_uniform_f32_device": (c_int, [c_int64, c_uint64, c_uint64, c_double,
                                                c_double, c_void_p]),
    "taco_hip_device_exclusive_scan": (c_int, [c_void_p, c_int32, c_void_p]),
    "taco_hip_csr_stats": (c_int, [P_T, POINTER(c_int64)]),
    "taco_hip_features_from_stats": (c_int, [c_void_p, c_int32, c_int32, c_void_p]),
    "taco_hip_features": (c_int, [P_T, c_int32, c_int32, c_void_p]),
    "taco_hip_mlp_param_count": (c_int64, []),
    "taco_hip_mlp_create": (c_int, [POINTER(c_void_p)]),
    "taco_hip_mlp_destroy": (None, [c_void_p]),
    "taco_hip_mlp_set_params": (c_int, [c_void_p, c_void_p]),
    "taco_hip_mlp_get_params": (c_int, [c_void_p, c_void_p]),
    "taco_hip_mlp_forward": (c_int, [c_void_p, c_void_p, c_int32, c_void_p]),
    "taco_hip_mlp_gradients": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                       c_void_p, POINTER(c_float)]),
    "taco_hip_mlp_train_step": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int32,
                                        c_float, c_float, POINTER(c_float)]),
    "taco_hip_workspace_get": (c_void_p, [c_size_t, c_int]),
    "taco_hip_generate": (c_void_p, [c_void_p, c_int32]),
    "taco_hip_free_string": (None, [c_void_p]),
    "taco_hip_module_create": (c_void_p, []),
    "taco_hip_module_set_source": (c_int, [c_void_p, c_char_p]),
    "taco_hip_module_get_source": (c_char_p, [c_void_p]),
    "taco_hip_module_compile": (c_int, [c_void_p]),
    "taco_hip_module_path": (c_char_p, [c_void_p]),
    "taco_hip_module_get_func": (c_void_p, [c_void_p, c_char_p]),
    "taco_hip_module_call_packed": (c_int, [c_void_p, c_char_p, c_void_p]),
    "taco_hip_module_destroy": (None, [c_void_p]),
    "taco_hip_module_report": (c_int, [c_int, c_char_p]),
    "taco_hip_module_finish": (c_int, [c_int]),
    "taco_hip_host_read_bytes": (c_int, [c_void_p, c_void_p, c_size_t]),
    "taco_hip_read_tensor_file": (c_int, [c_char_p, POINTER(c_void_p)]),
    "taco_hip_coo_info": (c_int, [c_void_p, POINTER(c_int32), c_void_p, c_int32,
                                  POINTER(c_int64), POINTER(c_int32)]),
    "taco_hip_coo_data": (c_int, [c_void_p, c_void_p, c_void_p]),
    "taco_hip_coo_destroy": (None, [c_void_p]),
    "taco_hip_write_tensor_file": (c_int, [c_char_p, c_int32, c_void_p, c_int64, c_void_p,
                                           c_void_p, c_int32]),
    "taco_hip_gather_ceiling": (c_int, [c_int32, c_int32, c_int32, c_int32, c_int64, c_int32,
                                        c_void_p]),
}


class GenSpec(Structure):
    """Mirror of ``taco_hip_gen_spec``."""
    _fields_ = [
        ("op", c_int32),
        ("family", c_int32),
        ("csize", c_int32),
        ("block_size", c_int32),
        ("nnz_per_thread", c_int32),
        ("warp_size", c_int32),
        ("rows_per_warp", c_int32),
        ("bound", c_int32),
    ]


OP_SPMV, OP_SPMM, OP_SPGEMM = range(3)
GEN_DEFAULT, GEN_WARP_ROWS, GEN_NNZ_SPLIT, GEN_TWO_PHASE, GEN_MERGE, GEN_BLOCK_ROWS = range(6)


def generate(spec, what=0):
    """Lowered IR (what=1) or HIP source (what=0) of a GenSpec."""
    L = lib()
    p = L.taco_hip_generate(ctypes.byref(spec), what)
    if not p:
        raise TacoHipError("generate failed: " +
                           L.taco_hip_last_error().decode(errors="replace"))
    try:
        return ctypes.string_at(p).decode()
    finally:
        L.taco_hip_free_string(p)


class Module:
    """JIT module: source -> hipcc -> .so (content-hash cached) -> dlopen."""

    def __init__(self, source):
        L = lib()
        self._m = L.taco_hip_module_create()
        check(L.taco_hip_module_set_source(self._m, source.encode()), "set_source")

    def compile(self):
        check(lib().taco_hip_module_compile(self._m), "module compile")
        return lib().taco_hip_module_path(self._m).decode()

    def call(self, name, descs):
        args = (c_void_p * len(descs))(*[ctypes.addressof(d.t) for d in descs])
        return lib().taco_hip_module_call_packed(self._m, name.encode(), args)

    def has(self, name):
        return bool(lib().taco_hip_module_get_func(self._m, name.encode()))

    def __del__(self):
        if getattr(self, "_m", None) and _lib is not None:
            _lib.taco_hip_module_destroy(self._m)
            self._m = None

EXPORTED = tuple(_SIGS)

_lib = None


class TacoHipError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it was not built or does not load."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TacoHipError(
            f"{LIB_PATH} is missing: build it with make -C taco_amd/csrc")
    handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def check(rc, what="taco_hip call"):
    if rc != 0:
        msg = lib().taco_hip_last_error().decode(errors="replace")
        raise TacoHipError(f"{what} failed ({rc}): {msg}")
    return rc


class SpmvPlan:
    """Slab-tiled SpMV plan over a CSR descriptor (see include/taco_hip.h)."""

    def __init__(self, A_desc, parts=0, slab_width=0):
        L = lib()
        self._p = c_void_p()
        check(L.taco_hip_spmv_plan_create(A_desc.ptr, parts, slab_width,
                                          ctypes.byref(self._p)), "spmv plan")
        self._keep = A_desc

    def info(self):
        out = (c_int64 * 5)()
        check(lib().taco_hip_spmv_plan_info(self._p, out))
        return dict(zip(("parts", "slabs", "slab_width", "row_weight",
                         "max_local_rows"), list(out)))

    def configure(self, threads=0, form=2):
        """Workgroup size (256 or 512) and kernel form: 2 interleaved entries
        (the default), 0 consecutive entries per thread, 1 wave-summed runs."""
        check(lib().taco_hip_spmv_plan_configure(self._p, threads, form), "plan configure")

    def shape(self):
        """The main kernel's launch shape: threads, form, default unroll."""
        t, f, u = c_int32(), c_int32(), c_int32()
        check(lib().taco_hip_spmv_plan_shape(self._p, ctypes.byref(t), ctypes.byref(f),
                                             ctypes.byref(u)), "plan shape")
        return {"threads": t.value, "form": f.value, "unroll": u.value}

    def time_kernel(self, launches):
        """Bracket the main kernel of the next `launches` computes with events."""
        check(lib().taco_hip_spmv_plan_time_kernel(self._p, launches), "plan timing")

    def kernel_time(self):
        """(mean main-kernel ms, recorded launches)."""
        ms, n = c_double(), c_int64()
        check(lib().taco_hip_spmv_plan_kernel_time(self._p, ctypes.byref(ms), ctypes.byref(n)))
        return (ms.value / n.value if n.value else float("nan")), n.value

    def compute(self, y_desc, x_desc, unroll=0):
        check(lib().taco_hip_spmv_plan_compute(self._p, y_desc.ptr, x_desc.ptr,
                                               unroll), "spmv plan compute")

    def __del__(self):
        if getattr(self, "_p", None) and _lib is not None:
            _lib.taco_hip_spmv_plan_destroy(self._p)
            self._p = None
