"""Tensor files (taco_amd.io over taco_hip_read_tensor_file): the reference's
own IO test files (copied from test/data into tests/golden/io) read to the
entries its tests expect (test/tests-io.cpp), symmetric expansion, dense
arrays with the first mode fastest, write/read round trips, and taco's
errors for headers it does not read."""
import os

import numpy as np
import pytest

from taco_amd import _lib, api as taco, io

IO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "io")


def entries(c):
    coords, vals = c.packed()
    return {tuple(int(v) for v in k): float(x) for k, x in zip(coords, vals)}


def test_tns_3tensor():
    c = io.read_coordinates(os.path.join(IO, "3tensor.tns"))
    assert c.dims == [1073, 1, 7]
    assert entries(c) == {(735, 0, 0): 1.0, (1072, 0, 5): 1.1, (880, 0, 6): 1.0}


def test_mtx_2tensor_with_blank_and_indented_comments():
    c = io.read_coordinates(os.path.join(IO, "2tensor.mtx"))
    assert c.dims == [32, 32]
    assert entries(c) == {(0, 0): 101.0, (1, 0): 102.0, (5, 2): 307.1}


def test_ttx_dense_array_first_mode_fastest():
    """The file lists 32 values for 4 x 3 x 2 positions: the first 24 fill
    the tensor, the rest are ignored (taco inserts `size` values)."""
    c = io.read_coordinates(os.path.join(IO, "d432.ttx"))
    assert c.dims == [4, 3, 2] and c.dense
    e = entries(c)
    v = 1.0
    for kk in range(2):
        for jj in range(3):
            for ii in range(4):
                assert e[(ii, jj, kk)] == v
                v += 1.0


def test_ttx_sparse_tensor():
    c = io.read_coordinates(os.path.join(IO, "d567.ttx"))
    assert c.dims == [5, 6, 7]
    assert entries(c) == {(0, 0, 0): 1.0, (1, 2, 0): 2.0, (4, 0, 6): 3.0, (2, 5, 2): 4.0}


def test_mtx_symmetric_expands():
    c = io.read_coordinates(os.path.join(IO, "ds33.mtx"))
    assert c.dims == [3, 3]
    assert entries(c) == {(0, 1): 1.0, (0, 2): 3.0, (1, 0): 1.0, (1, 1): 2.0, (2, 0): 3.0}


def test_coordinate_and_array_forms():
    """d33.mtx (coordinate) and d33.ttx (array) list the same values; read
    with the first mode fastest, as taco reads arrays, the array form is the
    transpose of the coordinate form."""
    a = taco.read(os.path.join(IO, "d33.mtx"))
    b = taco.read(os.path.join(IO, "d33.ttx"), taco.dense_matrix)
    pos, crd, vals = a.to_csr()
    dense = np.zeros((3, 3))
    for r in range(3):
        dense[r, crd[pos[r]:pos[r + 1]]] = vals[pos[r]:pos[r + 1]]
    assert np.array_equal(dense.T, b.to_numpy())
    assert dense.tolist() == [[0, 2, 0], [0, 0, 0], [3, 0, 4]]


def test_read_csr_matches_insert_and_pack():
    A = taco.read(os.path.join(IO, "rua_32.mtx"), taco.CSR, "A")
    c = io.read_coordinates(os.path.join(IO, "rua_32.mtx"))
    B = taco.Tensor("B", c.dims, taco.CSR)
    for k, v in zip(c.coords, c.vals):
        B.insert(list(k), float(v))
    B.pack()
    assert taco.equals(A, B)
    pos, crd, vals = A.to_csr()
    assert A.dims == [32, 32] and pos[-1] == len(c.vals) == 126


def test_repeated_coordinates_are_summed(tmp_path):
    """taco's pack sums repeated coordinates (test/tests-tensor.cpp:68-79);
    read() inserts every file entry the same way."""
    p = tmp_path / "dup.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n5 5 3\n2 3 42\n3 3 10\n2 3 1\n")
    A = taco.read(str(p))
    pos, crd, vals = A.to_csr()
    assert pos.tolist() == [0, 0, 1, 2, 2, 2] and crd.tolist() == [2, 2]
    assert vals.tolist() == [43.0, 10.0]
    T = taco.Tensor("T", [5, 5], taco.CSR)
    T.insert([1, 2], 42.0)
    T.insert([2, 2], 10.0)
    T.insert([1, 2], 1.0)
    T.pack()
    assert taco.equals(A, T)


@pytest.mark.parametrize("ext", ["mtx", "tns"])
def test_round_trip_is_exact(tmp_path, ext):
    from taco_amd import workloads as W
    Ah = W.csr_powerlaw(300, 200, seed=5, kmax=50)
    A = taco.Tensor("A", [300, 200], taco.CSR)
    A.set_csr(Ah.pos, Ah.crd, Ah.vals)
    p = str(tmp_path / f"a.{ext}")
    taco.write(p, A)
    B = taco.read(p)
    if ext == "tns":  # FROSTT dims are the largest coordinates seen
        assert B.dims[0] <= 300 and B.dims[1] <= 200
        B = taco.read(p)
    pa, ca, va = A.to_csr()
    pb, cb, vb = B.to_csr()
    assert np.array_equal(ca, cb) and np.array_equal(va, vb)
    assert np.array_equal(pa[:len(pb)], pb)


def test_dense_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.uniform(-1, 1, (4, 7))
    D = taco.Tensor("D", [4, 7], taco.dense_matrix)
    D.set_values(a)
    p = str(tmp_path / "d.mtx")
    taco.write(p, D)
    assert open(p).readline().strip() == "%%MatrixMarket matrix array real general"
    assert np.array_equal(taco.read(p, taco.dense_matrix).to_numpy(), a)


@pytest.mark.parametrize("header,msg", [
    ("%%MatrixMarket matrix coordinate integer general", "field not available"),
    ("%%MatrixMarket matrix coordinate real hermitian", "symmetry not available"),
    ("%%MatrixMarket vector coordinate real general", "unknown type"),
    ("%%Matrix matrix coordinate real general", "unknown header"),
])
def test_headers_taco_does_not_read_are_errors(tmp_path, header, msg):
    p = tmp_path / "bad.mtx"
    p.write_text(header + "\n2 2 1\n1 1 1\n")
    with pytest.raises(_lib.TacoHipError, match=msg):
        io.read_coordinates(str(p))


def test_out_of_range_and_short_files_are_errors(tmp_path):
    p = tmp_path / "oob.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n")
    with pytest.raises(_lib.TacoHipError, match="out of range"):
        io.read_coordinates(str(p))
    p = tmp_path / "short.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n")
    with pytest.raises(_lib.TacoHipError, match="announces 2 entries"):
        io.read_coordinates(str(p))
    with pytest.raises(_lib.TacoHipError, match="not supported"):
        io.read_coordinates(str(tmp_path / "x.rb"))
