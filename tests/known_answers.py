"""Known-answer cases transcribed from the reference's tests (inputs and
expected outputs are literal values there).

  spmv_d33a_d3b     test/tests-expr_storage.cpp:873-937 with fixtures
                    test/test_tensors.cpp:89-94 (d3b) and :205-211 (d33a):
                    a(i) = B(i,k) * c(k) -> {0, 0, 18}
  matrix_vector_mul test/tests-lower.cpp:612-632, 5x5 default shapes
                    (tests-lower.cpp:76-84): a = {0: 2, 3: 8}
  matrix_mul        test/tests-expr_storage.cpp:996-1035 with d33a, d33b
                    (test_tensors.cpp:221-227): [[0,0,0],[0,0,0],[30,180,0]]
  where_spmm        test/tests-lower.cpp:736-754, CSR result on 5x5:
                    A = {(2,0): 30, (2,1): 180}, pos {0,0,0,2,2,2}, crd {0,1}
"""
import numpy as np


def coo_to_csr(m, n, entries):
    entries = sorted(entries)
    pos = np.zeros(m + 1, dtype=np.int32)
    crd, vals = [], []
    for (i, j), v in entries:
        pos[i + 1] += 1
        crd.append(j)
        vals.append(v)
    pos = np.cumsum(pos).astype(np.int32)
    return pos, np.array(crd, dtype=np.int32), np.array(vals, dtype=np.float64)


D33A = [((0, 1), 2.0), ((2, 0), 3.0), ((2, 2), 4.0)]
D33B = [((0, 0), 10.0), ((0, 1), 20.0), ((2, 1), 30.0)]
D3B = {0: 2.0, 2: 3.0}

MVM_B = [((0, 0), 42.0), ((0, 2), 2.0), ((1, 1), 3.0), ((3, 3), 4.0)]
MVM_C = {2: 1.0, 3: 2.0}
MVM_EXPECTED = {0: 2.0, 3: 8.0}

SPMM_B = [((0, 1), 2.0), ((2, 0), 3.0), ((2, 2), 4.0)]
SPMM_C = [((0, 0), 10.0), ((0, 1), 20.0), ((2, 1), 30.0)]
SPMM_EXPECTED = {(2, 0): 30.0, (2, 1): 180.0}
SPMM_EXPECTED_POS = [0, 0, 0, 2, 2, 2]
SPMM_EXPECTED_CRD = [0, 1]
