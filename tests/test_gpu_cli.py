"""taco-hip -hip -time on the GPU: operands read from files, taco's lowering
with CodeGen_HIP, assembled and computed through taco's TensorBase, timed
with HIP events; the written result is checked against the FP64 oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from taco_amd import api as taco, io
from taco_amd import workloads as W

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "taco_amd", "upstream", "taco-hip")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(CLI), reason="taco-hip not built")]


def write_operands(tmp_path, A, x):
    At = taco.Tensor("A", [A.m, A.n], taco.CSR)
    At.set_csr(A.pos, A.crd, A.vals)
    taco.write(str(tmp_path / "A.mtx"), At)
    xt = taco.Tensor("x", [A.n], taco.dense_vector)
    xt.set_values(x)
    taco.write(str(tmp_path / "x.mtx"), xt)


@pytest.mark.parametrize("schedule", [None, "split(i,i0,i1,256),parallelize(i0,GPUBlock,NoRaces),"
                                            "parallelize(i1,GPUThread,NoRaces)"])
def test_cli_spmv_times_and_matches_oracle(tmp_path, schedule):
    A = W.csr_powerlaw(20000, 15000, seed=81, kmax=2000)
    x = W.uniform(A.n, 81, W.STREAM_X)
    write_operands(tmp_path, A, x)
    args = [CLI, "y(i) = A(i,j) * x(j)", "-f=A:ds", "-f=x:d", "-f=y:d", "-i=A:A.mtx",
            "-i=x:x.mtx", "-o=y:y.mtx", "-hip", "-time=5"]
    if schedule:
        args.append("-s=" + schedule)
    print("taco-hip", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"Compute \(HIP events\):\s+mean:\s+([0-9.]+) ms", r.stdout)
    assert m and float(m.group(1)) > 0, r.stdout
    y = io.read_coordinates(str(tmp_path / "y.mtx"))
    assert y.dims == [A.m]
    got = np.zeros(A.m)
    got[y.coords[:, 0]] = y.vals
    ye, cond = oracle.exact_spmv(A, x)
    rep = oracle.error_report(got, ye, cond, 1e-12)
    assert rep["ok"], rep
