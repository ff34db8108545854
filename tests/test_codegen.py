"""CodeGen_HIP and the JIT module loader, on the CPU (no GPU calls): the
lowered IR carries the schedule's parallel units, the emitted module has the
taco module ABI, modules compile through the cache, and failures are
reported, not fatal."""
import os
import subprocess
import sys

import pytest

from taco_amd import _lib, codegen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("name", sorted(codegen.STANDARD_SPECS))
def test_module_abi(name):
    src = _lib.generate(codegen.standard(name))
    for fn in ("assemble", "compute", "evaluate"):
        assert f'extern "C" int {fn}(taco_tensor_t*' in src
        assert f'extern "C" int _shim_{fn}(void** parameterPack)' in src
    assert "__global__ __launch_bounds__(" in src
    assert "hipLaunchKernelGGL(computeDeviceKernel0" in src
    # no device-wide sync after launches, no CUDA spellings
    for bad in ("cudaDeviceSynchronize", "hipDeviceSynchronize", "__syncwarp",
                "__shfl_sync", "cudaMallocManaged", "hipMemsetAsync"):
        assert bad not in src, bad


def test_ir_default_schedule_units():
    ir = _lib.generate(codegen.standard("spmv_default_f64"), 1)
    assert "for i0 in [0, ((A1_dimension + 255) / 256)) step 1 GPUBlock" in ir
    assert "for i1 in [0, 256) step 1 GPUThread" in ir


def test_ir_warp_rows_units():
    ir = _lib.generate(codegen.standard("spmv_rows_f64"), 1)
    for unit in ("GPUBlock", "GPUWarp", "GPUThread", "GPUWarpReduction"):
        assert unit in ir
    src = _lib.generate(codegen.standard("spmv_rows_f64"))
    assert "taco_hip_group_sum<32>" in src
    assert "threadIdx.x / 32" in src and "threadIdx.x % 32" in src


def test_ir_spmm_warp_rows():
    """scheduleSpMMRowsGPU: 8 warps of R = 4 rows per 256-thread block, k in
    32-column chunks over the warp's lanes, atomic adds into C."""
    ir = _lib.generate(codegen.standard("spmm_rows_f64"), 1)
    for unit in ("GPUBlock", "GPUWarp", "GPUThread"):
        assert unit in ir
    src = _lib.generate(codegen.standard("spmm_rows_f64"))
    assert "((block * 32) + ((warp * 4) + warp_row))" in src
    assert "((B2_dimension + 31) / 32)" in src
    assert "taco_hip_atomicAdd(C_vals" in src and "taco_hip_fill(C_vals" in src


def test_ir_nnz_split_atomics():
    ir = _lib.generate(codegen.standard("spmv_nnz_f64"), 1)
    assert "[atomic GPUThread]" in ir and "[atomic GPUWarp]" in ir
    src = _lib.generate(codegen.standard("spmv_nnz_f64"))
    assert "taco_hip_binarySearchBeforeBlockLaunch" in src
    assert "taco_hip_atomicAddWave(y_vals" in src
    assert "#pragma unroll 8" in src


def test_ir_block_rows_reduction():
    ir = _lib.generate(codegen.standard("spmv_block_f64"), 1)
    for unit in ("GPUBlock", "GPUThread", "GPUBlockReduction"):
        assert unit in ir
    assert "GPUWarp " not in ir
    src = _lib.generate(codegen.standard("spmv_block_f64"))
    assert "taco_hip_block_sum<256>(tthread_val)" in src
    assert "if (thread == 0) { y_vals[i] = taco_red; }" in src
    for bad_block in (96, 2048, 48):
        with pytest.raises(_lib.TacoHipError, match="BLOCK_ROWS"):
            _lib.generate(codegen.spec(_lib.OP_SPMV, _lib.GEN_BLOCK_ROWS, block_size=bad_block))
    with pytest.raises(_lib.TacoHipError):
        _lib.generate(codegen.spec(_lib.OP_SPMM, _lib.GEN_BLOCK_ROWS))


def test_nnz_split_full_chunk_paths():
    """Full chunks skip the per-entry guard and stage their entries (and, for
    SpMM, their B values) in registers; partial chunks keep the guard."""
    src = _lib.generate(codegen.standard("spmv_nnz_f64"))
    assert "if ((fpos0 <= (A_nnz - 8)))" in src
    assert "chunk_crd[thread_nz] = A2_crd[(fpos0 + thread_nz)];" in src
    assert "chunk_x[thread_nz] = x_vals[chunk_crd[thread_nz]];" in src
    assert "if ((fposA < A_nnz))" in src  # the guarded partial chunk
    src = _lib.generate(codegen.standard("spmm_nnz_f64"))
    assert "double chunk_b[8];" in src and "if ((fpos0 <= (A_nnz - 8)))" in src


def test_merge_spmm_batches_gathers():
    src = _lib.generate(codegen.standard("spmm_merge_f64"))
    assert "for (int32_t u0 = 0; u0 < (cnt - 7); u0 += 8)" in src
    assert "batch_b[q] = B_vals[" in src
    assert "for (int32_t u = ((cnt / 8) * 8); u < cnt; u += 1)" in src  # tail


def test_spgemm_compute_strides_rows():
    src = _lib.generate(codegen.standard("spgemm_f64"))
    assert "int32_t row_blocks = taco_hip_row_grid(A1_dimension);" in src
    assert "i += (row_blocks * 4)" in src
    # assemble keeps one workgroup per four rows
    assert src.count("taco_hip_row_grid") == 1


def test_kernel_parameters_are_values_not_references():
    src = _lib.generate(codegen.standard("spmm_nnz_f64"))
    kernel_sig = [l for l in src.splitlines() if l.startswith("void computeDeviceKernel0(")][0]
    assert "&" not in kernel_sig
    assert "int32_t A_nnz" in kernel_sig


def test_unsupported_spec_reports_error():
    bad = codegen.spec(_lib.OP_SPGEMM, _lib.GEN_NNZ_SPLIT)
    with pytest.raises(_lib.TacoHipError, match="two-phase"):
        _lib.generate(bad)
    with pytest.raises(_lib.TacoHipError, match="f64 and f32"):
        _lib.generate(codegen.spec(_lib.OP_SPMV, 0, csize=16))


def test_module_compiles_and_caches():
    k1 = codegen.Kernel(codegen.standard("spmv_default_f64"))
    assert os.path.exists(k1.path) and k1.module.has("_shim_compute")
    k2 = codegen.Kernel(codegen.standard("spmv_default_f64"))
    assert k2.path == k1.path


def test_compile_failure_is_reported(tmp_path):
    m = _lib.Module("this is not HIP code;\n")
    with pytest.raises(_lib.TacoHipError, match="failed"):
        m.compile()


def test_missing_compiler_is_reported(tmp_path):
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "from taco_amd import _lib\n"
        "m = _lib.Module('// fresh module %s\\n')\n"
        "try:\n    m.compile()\nexcept _lib.TacoHipError as e:\n    print('ERR', e)\n"
    ) % (ROOT, os.getpid())
    env = dict(os.environ, TACO_HIPCC="/nonexistent/hipcc", TACO_HIP_CACHE=str(tmp_path))
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True,
                         text=True, timeout=120)
    assert "ERR" in out.stdout and "/nonexistent/hipcc" in out.stdout
