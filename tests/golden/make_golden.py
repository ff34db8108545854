"""Writes the golden fixtures of tests/golden from the real taco.

Run in a checkout where oracle/_ref/libtaco_ref.so was built from the
reference sources (`make -C oracle ref`):

    python tests/golden/make_golden.py

Fixtures (data only: inputs and taco's outputs):
  ref_tests.npz   the inputs of the reference's own GPU/CPU schedule tests,
                  regenerated exactly (glibc srand/rand, float arithmetic as
                  in test/tests-scheduling-eval.cpp), with the outputs of
                  taco's default CPU schedule:
                    spmvGPU  (:1210-1256)  102x103, 1% dense, srand(94353)
                    spmmGPU  (:1258-1307)  102x103 x 128, 30%, srand(434321)
                    spmvCPU  (:843-889)    102x103, 30%, srand(120)
  seeded.npz      seeded random CSR operands with U[-1,1] values and taco's
                  outputs for SpMV, SpMM (K=16) and two-phase SpGEMM.
"""
import ctypes
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from taco_amd import workloads as W  # noqa: E402

libc = ctypes.CDLL("libc.so.6")
RAND_MAX = 2147483647


def rand_float():
    # float rand_float = (float)rand()/(float)(RAND_MAX);
    return np.float32(libc.rand()) / np.float32(RAND_MAX)


def tval(rf, sparsity):
    # (double)((int)(rand_float * 3 / SPARSITY)), float arithmetic
    return float(int(np.float32(np.float32(rf * np.float32(3)) / np.float32(sparsity))))


def sparse_from_rand(ni, nj, sparsity):
    rows = [[] for _ in range(ni)]
    for i in range(ni):
        for j in range(nj):
            rf = rand_float()
            if rf < np.float32(sparsity):
                rows[i].append((j, tval(rf, sparsity)))
    pos = np.zeros(ni + 1, dtype=np.int32)
    crd, vals = [], []
    for i in range(ni):
        for j, v in rows[i]:
            crd.append(j)
            vals.append(v)
        pos[i + 1] = len(crd)
    return W.Csr(ni, nj, pos, np.array(crd, dtype=np.int32),
                 np.array(vals, dtype=np.float64))


def ref_tests():
    out = {}
    # spmvGPU
    ni, nj, sp = 1021 // 10, 1039 // 10, 0.01
    libc.srand(94353)
    A = sparse_from_rand(ni, nj, sp)
    x = np.array([tval(rand_float(), sp) for _ in range(nj)])
    y = oracle.ref_spmv(A, x)
    out.update(spmvgpu_pos=A.pos, spmvgpu_crd=A.crd, spmvgpu_vals=A.vals,
               spmvgpu_x=x, spmvgpu_y=y)
    # spmmGPU
    nk, sp = 128, 0.3
    libc.srand(434321)
    A = sparse_from_rand(ni, nj, sp)
    B = np.array([[tval(rand_float(), sp) for _ in range(nk)] for _ in range(nj)])
    C = oracle.ref_spmm(A, B)
    out.update(spmmgpu_pos=A.pos, spmmgpu_crd=A.crd, spmmgpu_vals=A.vals,
               spmmgpu_B=B, spmmgpu_C=C)
    # spmvCPU
    libc.srand(120)
    A = sparse_from_rand(ni, nj, sp)
    x = np.array([tval(rand_float(), sp) for _ in range(nj)])
    y = oracle.ref_spmv(A, x)
    out.update(spmvcpu_pos=A.pos, spmvcpu_crd=A.crd, spmvcpu_vals=A.vals,
               spmvcpu_x=x, spmvcpu_y=y)
    return out


def seeded():
    out = {}
    A = W.csr_powerlaw(1500, 1200, seed=41, kmax=1000)
    x = W.uniform(A.n, 41, W.STREAM_X)
    out.update(spmv_pos=A.pos, spmv_crd=A.crd, spmv_vals=A.vals, spmv_x=x,
               spmv_y=oracle.ref_spmv(A, x, threads=4))
    B = W.uniform(A.n * 16, 42, W.STREAM_B).reshape(A.n, 16)
    out.update(spmm_B=B, spmm_C=oracle.ref_spmm(A, B, threads=4))
    G1 = W.csr_fixed(800, 600, 8, 43)
    G2 = W.csr_powerlaw(600, 700, seed=44, kmax=300)
    p, c, v = oracle.ref_spgemm(G1, G2, threads=4)
    out.update(g_apos=G1.pos, g_acrd=G1.crd, g_avals=G1.vals,
               g_bpos=G2.pos, g_bcrd=G2.crd, g_bvals=G2.vals,
               g_bn=np.array([G2.n]), g_cpos=p, g_ccrd=c, g_cvals=v)
    return out


def main():
    if not oracle.ref_available():
        sys.exit("oracle/_ref/libtaco_ref.so not built (make -C oracle ref)")
    here = os.path.dirname(os.path.abspath(__file__))
    np.savez_compressed(os.path.join(here, "ref_tests.npz"), **ref_tests())
    np.savez_compressed(os.path.join(here, "seeded.npz"), **seeded())
    print("wrote", os.listdir(here))


if __name__ == "__main__":
    main()
