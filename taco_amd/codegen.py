"""Python face of CodeGen_HIP and the JIT module loader.

``Kernel(spec)`` lowers a statement under a schedule family, prints it with
CodeGen_HIP, compiles it through the module cache and exposes the module's
assemble/compute/evaluate functions, called with taco_tensor_t descriptors
in taco's argument order (results first, then operands).
"""
from . import _lib

# Named schedules: the reference's GPU schedules at their default parameters
# (test/tests-scheduling-eval.cpp), the f32 forms, and the HIP backend's
# merge-path SpMV and SpMM, and the block-per-row SpMV (GPUBlockReduction).
# spmm_rows_f64 is scheduleSpMMRowsGPU (4 rows per warp, 32-column chunks);
# spmm_merge16/32 are merge-path SpMM with teams of 16 / 32 lanes (K <= 128).
STANDARD_SPECS = {
    "spmv_default_f64": (_lib.OP_SPMV, _lib.GEN_DEFAULT, 64, 256, 0, 0, 0, 0),
    "spmv_rows_f64": (_lib.OP_SPMV, _lib.GEN_WARP_ROWS, 64, 256, 0, 32, 1, 0),
    "spmv_rows64_f64": (_lib.OP_SPMV, _lib.GEN_WARP_ROWS, 64, 256, 0, 64, 1, 0),
    "spmv_nnz_f64": (_lib.OP_SPMV, _lib.GEN_NNZ_SPLIT, 64, 256, 8, 32, 0, 0),
    "spmv_merge_f64": (_lib.OP_SPMV, _lib.GEN_MERGE, 64, 256, 8, 0, 0, 0),
    "spmv_block_f64": (_lib.OP_SPMV, _lib.GEN_BLOCK_ROWS, 64, 256, 0, 0, 0, 0),
    "spmv_default_f32": (_lib.OP_SPMV, _lib.GEN_DEFAULT, 32, 256, 0, 0, 0, 0),
    "spmv_nnz_f32": (_lib.OP_SPMV, _lib.GEN_NNZ_SPLIT, 32, 256, 8, 32, 0, 0),
    "spmm_default_f64": (_lib.OP_SPMM, _lib.GEN_DEFAULT, 64, 256, 0, 0, 0, 0),
    "spmm_nnz_f64": (_lib.OP_SPMM, _lib.GEN_NNZ_SPLIT, 64, 256, 8, 32, 0, 4),
    "spmm_nnz_f32": (_lib.OP_SPMM, _lib.GEN_NNZ_SPLIT, 32, 256, 8, 32, 0, 4),
    "spmm_merge_f64": (_lib.OP_SPMM, _lib.GEN_MERGE, 64, 256, 256, 64, 0, 2),
    "spmm_rows_f64": (_lib.OP_SPMM, _lib.GEN_WARP_ROWS, 64, 256, 0, 32, 4, 0),
    "spmm_merge16_f64": (_lib.OP_SPMM, _lib.GEN_MERGE, 64, 256, 256, 16, 0, 8),
    "spmm_merge32_f64": (_lib.OP_SPMM, _lib.GEN_MERGE, 64, 256, 256, 32, 0, 4),
    "spgemm_f64": (_lib.OP_SPGEMM, _lib.GEN_TWO_PHASE, 64, 256, 0, 0, 0, 0),
}


# Modules the unified-memory tests compile with unified memory on (their
# flags, and so their cache entries, differ: compare-and-swap atomics).
UNIFIED_PREBUILT = ("spmv_default_f64", "spmv_nnz_f64", "spmv_merge_f64",
                    "spmm_nnz_f64", "spmm_merge_f64")


def spec(op, family, csize=64, block_size=256, nnz_per_thread=0, warp_size=0,
         rows_per_warp=0, bound=0):
    return _lib.GenSpec(op, family, csize, block_size, nnz_per_thread,
                        warp_size, rows_per_warp, bound)


def standard(name):
    return spec(*STANDARD_SPECS[name])


class Kernel:
    def __init__(self, gen_spec):
        self.spec = gen_spec
        self.source = _lib.generate(gen_spec, 0)
        self.ir = _lib.generate(gen_spec, 1)
        self.module = _lib.Module(self.source)
        self.path = self.module.compile()

    def _call(self, name, descs):
        rc = self.module.call(name, descs)
        _lib.check(rc, f"generated {name}")

    def assemble(self, *descs):
        self._call("assemble", descs)

    def compute(self, *descs):
        self._call("compute", descs)

    def evaluate(self, *descs):
        self._call("evaluate", descs)


def prebuild_cache(names=None):
    """Compiles the standard modules into the JIT cache (no GPU needed),
    and the UNIFIED_PREBUILT ones also as compiled under unified memory."""
    paths = {}
    for name in names or STANDARD_SPECS:
        paths[name] = Kernel(standard(name)).path
    L = _lib.lib()
    was = L.taco_hip_should_use_unified_memory()
    L.taco_hip_set_unified_memory_enabled(1)
    try:
        for name in UNIFIED_PREBUILT:
            paths[name + "@unified"] = Kernel(standard(name)).path
    finally:
        L.taco_hip_set_unified_memory_enabled(was)
    return paths
