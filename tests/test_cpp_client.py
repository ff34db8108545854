"""A C++ program calling the C ABI the way taco's TensorBase would
(tests/cpp/taco_client.cpp): descriptors, CodeGen_HIP source, the module
loader, _shim_ calls with taco's argument order and output ownership."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "taco_amd")


def build_client(tmp_path):
    exe = tmp_path / "taco_client"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", "taco_client.cpp"), "-L", LIBDIR,
                    "-ltaco_hip", f"-Wl,-rpath,{LIBDIR}", "-o", str(exe)], check=True)
    return exe


def test_cpp_client_generates_and_compiles(tmp_path):
    out = subprocess.run([str(build_client(tmp_path)), "generate"], capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr
    assert "generate ok" in out.stdout and out.stdout.count(".so") == 2


@pytest.mark.gpu
def test_cpp_client_runs_on_the_gpu(tmp_path):
    print("cpp client: spmv (nnz split) and two-phase spgemm modules", flush=True)
    out = subprocess.run([str(build_client(tmp_path)), "run"], capture_output=True, text=True,
                         timeout=300)
    print(out.stdout, out.stderr, flush=True)
    assert out.returncode == 0, out.stderr
    assert "spmv ok" in out.stdout and "spgemm ok" in out.stdout and "run ok" in out.stdout
