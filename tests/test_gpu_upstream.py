"""The reference's GPU test statements (test/tests-scheduling-eval.cpp:
1210-1588, test/tests-scheduling.cpp:393-536, test/tests-windowing.cpp:
342-386) run through taco itself — its API, scheduling language, lowering
and TensorBase compile/assemble/compute — with CodeGen_HIP, the HIP Module
and the HIP runtime (unified memory) plugged in.  Inputs are the reference
tests' own rand() draws; each result is compared exactly with dense brute
force over the same small-integer values (two statements whose taco IR
stores rather than accumulates are checked for what that store leaves; see
taco_amd/upstream/drop_in_tests.cpp)."""
import pytest

from taco_amd import upstream

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not upstream.available(), reason="taco_amd/upstream not built")]

NAMES = ["spmv_rows_gpu", "spmv_splitpos_gpu", "spmv_thread_row_gpu", "spmv_default_gpu",
         "spmv_precompute_gpu", "spmm_gpu", "spmm_rows_gpu", "spmm_dcsr_gpu", "sddmm_gpu",
         "ttm_gpu", "ttv_gpu", "mttkrp_gpu", "atomic_reduction_gpu",
         "temporary_reduction_gpu", "tiling_gpu", "windowing_gpu_dd", "windowing_gpu_ds",
         "windowing_gpu_sd", "windowing_gpu_ss"]


@pytest.fixture(scope="module", autouse=True)
def library_defaults_after():
    # taco's runtime switches the library to unified memory and blocking
    # calls; later tests in this process expect the library defaults.
    yield
    from taco_amd import _lib
    _lib.lib().taco_hip_set_unified_memory_enabled(0)
    _lib.lib().taco_hip_set_blocking(0)


def test_statement_list_matches():
    assert upstream.names() == NAMES


@pytest.mark.parametrize("name", NAMES)
def test_reference_gpu_statement(name):
    print(f"drop-in statement {name}", flush=True)
    ok, msg = upstream.run(name)
    assert ok, f"{name}: {msg}"
