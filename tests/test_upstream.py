"""taco's own lowering with CodeGen_HIP plugged in (taco_amd/upstream): the
HIP module text printed for each of the reference's GPU test statements, on
the CPU.  The statements run on the GPU in tests/test_gpu_upstream.py."""
import re

import pytest

from taco_amd import upstream

pytestmark = pytest.mark.skipif(not upstream.available(),
                                reason="taco_amd/upstream not built (needs the reference tree)")


def test_all_reference_gpu_statements_print():
    names = upstream.names()
    assert len(names) == 19
    for n in names:
        src = upstream.emit(n)
        assert "int compute(" in src and "_shim_compute" in src, n
        # wave64 CDNA code: no CUDA warp primitives, no device-wide syncs
        for banned in ("__syncwarp", "__shfl_sync", "cudaDeviceSynchronize",
                       "cudaMallocManaged", "atomicAddWarp"):
            assert banned not in src, (n, banned)


def test_kernels_take_values_not_descriptors():
    """Every property is unpacked on the host: kernels take raw pointers and
    sizes (the reference passes taco_tensor_t* into kernels)."""
    src = upstream.emit("spmm_gpu")
    kernels = re.findall(r"__global__[^\n]*\nvoid (\w+)\(([^)]*)\)", src)
    assert kernels
    for name, params in kernels:
        assert "taco_tensor_t" not in params, name


def test_warp_reduction_is_wave_local():
    """scheduleSpMVRowsGPU's 32-lane groups sit inside one 64-lane wave:
    the group reduction needs wave-level ordering only."""
    src = upstream.emit("spmv_rows_gpu")
    assert "taco_hip_wave_sync();" in src
    assert "__launch_bounds__(256)" in src


def test_runtime_workgroup_shape():
    """scheduleSpMMRowsGPU's thread loop extent is ceil(K/32): the workgroup
    size is computed on the host and checked before the launch."""
    src = upstream.emit("spmm_rows_gpu")
    assert "taco_threads" in src and "workgroup size out of range" in src


def test_host_scalar_reduced_through_device_copy():
    """parallelizeAtomicReduction adds into a host scalar from every thread:
    a device copy, compare-and-swap atomics, value copied back."""
    src = upstream.emit("atomic_reduction_gpu")
    assert "taco_hip_copy_in(C_val_ref" in src and "taco_hip_copy_out(&C_val" in src
    assert "taco_hip_atomicAdd(C_val_ref, 0," in src


def test_names_do_not_depend_on_process_counters():
    """taco numbers unnamed tensors and index variables with a process-wide
    counter; the printed text must not, so that the module cache keyed by
    the text serves every process."""
    a = upstream.emit("spmv_default_gpu")
    for n in ("spmm_gpu", "ttm_gpu", "sddmm_gpu"):
        upstream.emit(n)  # advances taco's counters
    assert upstream.emit("spmv_default_gpu") == a
