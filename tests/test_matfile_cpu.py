"""Level 5 MAT-file input (include/graph.h: load_csr_from_mat,
load_csr_from_file; reference src/graph.c:14-128).

The reader is pinned against scipy.io.savemat files (an independent writer
of the format): plain and zlib-compressed variables, a dense variable before
the sparse one, a "Problem" struct whose field "A" is taken, rectangular and
valued matrices.  The expected CSR is the transpose of the stored CSC, i.e.
scipy's own CSR of the same matrix with sorted indices.  The writer is
pinned the other way: scipy.io.loadmat must read back the same matrix for
every flag combination (byte order, compression, struct wrapping)."""
import os

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as sp

import cc_mpi_amd as cm
from cc_mpi_amd import graphs
import oracle


def csr_of(M):
    C = sp.csr_matrix(M)
    C.sort_indices()
    return C.indptr.astype(np.uint64), C.indices.astype(np.uint32)


def assert_graph(G, M):
    rp, ci = csr_of(M)
    assert G.n == M.shape[0]
    np.testing.assert_array_equal(G.row_ptr, rp)
    np.testing.assert_array_equal(G.col_idx, ci)


@pytest.mark.parametrize("compress", [False, True])
def test_scipy_sparse_logical(tmp_path, compress):
    A = sp.random(300, 300, density=0.03, format="csc", random_state=1)
    B = ((A + A.T) != 0).astype(bool)
    p = str(tmp_path / "g.mat")
    sio.savemat(p, {"B": B}, do_compression=compress)
    assert_graph(cm.load(p), B)


@pytest.mark.parametrize("compress", [False, True])
def test_first_sparse_variable_after_dense(tmp_path, compress):
    """Dense variables are skipped; the first sparse one is taken."""
    A = sp.random(40, 40, density=0.2, format="csc", random_state=2)
    A2 = sp.random(40, 40, density=0.5, format="csc", random_state=3)
    p = str(tmp_path / "g.mat")
    sio.savemat(p, {"x": np.arange(7.0), "name": "text", "first": A, "second": A2},
                do_compression=compress)
    assert_graph(cm.load(p), A)


@pytest.mark.parametrize("compress", [False, True])
def test_problem_struct_field_A(tmp_path, compress):
    """A struct named Problem supplies its field A, wherever A sits among
    the fields, ahead of sparse variables stored before it."""
    A = sp.random(64, 64, density=0.1, format="csc", random_state=4)
    other = sp.random(64, 64, density=0.1, format="csc", random_state=5)
    p = str(tmp_path / "p.mat")
    sio.savemat(p, {"Problem": {"title": "t", "id": 7, "A": A, "notes": np.ones(3)}},
                do_compression=compress)
    assert_graph(cm.load(p), A)
    sio.savemat(p, {"aaa": np.zeros(2), "Problem": {"A": A}, "zz": other},
                do_compression=compress)
    assert_graph(cm.load(p), A)


def test_rows_are_vertices_rectangular(tmp_path):
    """n = number of rows; row r lists the columns with an entry in row r,
    column ids kept as stored (the reference does not check them)."""
    A = sp.random(50, 70, density=0.1, format="csc", random_state=6)
    p = str(tmp_path / "r.mat")
    sio.savemat(p, {"A": A})
    G = cm.load(p)
    assert_graph(G, A)
    assert G.col_idx.max() >= 50


def test_values_ignored(tmp_path):
    A = sp.random(30, 30, density=0.2, format="csc", random_state=7) * 1e6
    A.data[::3] = -2.5
    p = str(tmp_path / "v.mat")
    sio.savemat(p, {"A": A})
    assert_graph(cm.load(p), A)


def test_empty_and_isolated(tmp_path):
    p = str(tmp_path / "e.mat")
    sio.savemat(p, {"A": sp.csc_matrix((9, 9), dtype=bool)})
    G = cm.load(p)
    assert G.n == 9 and G.m == 0
    np.testing.assert_array_equal(G.row_ptr, np.zeros(10, np.uint64))


@pytest.mark.parametrize("flags", range(8))
@pytest.mark.parametrize("name", ["G", "Adj", "graph_er"])
def test_writer_read_by_scipy_and_back(tmp_path, flags, name):
    G = cm.erdos_renyi(2000, 6.0, seed=11)
    p = str(tmp_path / "w.mat")
    cm.save_mat(p, G, name, flags)
    d = sio.loadmat(p)
    M = d["Problem"]["A"][0, 0] if flags & graphs.MAT_PROBLEM else d[name]
    if flags & graphs.MAT_PROBLEM:
        assert str(d["Problem"]["name"][0, 0][0]) == name
    assert sp.issparse(M) and M.shape == (G.n, G.n)
    assert_graph(G, M)
    H = cm.load(p)
    np.testing.assert_array_equal(H.row_ptr, G.row_ptr)
    np.testing.assert_array_equal(H.col_idx, G.col_idx)


def test_non_symmetric_orientation(tmp_path):
    """A directed CSR survives write -> read unchanged (rows stay rows)."""
    from graphs_util import csr_raw
    G = csr_raw(5, [[1, 4], [], [0], [0, 1, 2], [3]])
    p = str(tmp_path / "d.mat")
    cm.save_mat(p, G, "D")
    M = sp.csr_matrix(sio.loadmat(p)["D"])
    assert M[3, 2] and not M[2, 3]
    H = cm.load(p)
    np.testing.assert_array_equal(H.row_ptr, G.row_ptr)
    np.testing.assert_array_equal(H.col_idx, G.col_idx)


def test_er64k_input_labels(tmp_path):
    """The reference's CPU-runnable case: ER 65,536 vertices, average degree
    8, symmetric and simple, stored as a sparse logical .mat; the graph read
    back is the graph written, so its labels are the same."""
    G = cm.erdos_renyi(65536, 8.0, seed=1001)
    p = str(tmp_path / "er64k.mat")
    cm.save_mat(p, G, "A", graphs.MAT_PROBLEM)
    H = cm.load(p)
    np.testing.assert_array_equal(H.row_ptr, G.row_ptr)
    np.testing.assert_array_equal(H.col_idx, G.col_idx)
    want = oracle.cc_min_label(G.row_ptr, G.col_idx)
    np.testing.assert_array_equal(oracle.cc_min_label(H.row_ptr, H.col_idx), want)


def test_error_codes(tmp_path):
    lib = graphs.nat.host_lib()
    g = graphs.nat.CSRGraphC()
    def rc(path):
        return lib.load_csr_from_file(str(path).encode(), 1, 1, graphs.ctypes.byref(g))
    assert rc(tmp_path / "missing.mat") == 1
    bad = tmp_path / "bad.mat"
    bad.write_bytes(b"not a mat file" * 20)
    assert rc(bad) == 1
    dense = tmp_path / "dense.mat"
    sio.savemat(str(dense), {"x": np.eye(3)})
    assert rc(dense) == 2
    noa = tmp_path / "noa.mat"
    sio.savemat(str(noa), {"Problem": {"B": sp.eye(3, format="csc")}, "S": sp.eye(3, format="csc")})
    assert rc(noa) == 2  # the Problem struct stops the search, as in the reference
    assert rc(tmp_path / "g.mtx") == -1
    assert rc(tmp_path / "g.txt") == -1
    assert rc(tmp_path / "g.bin") == 99
    assert rc(tmp_path / "noext") == 99
    with pytest.raises(graphs.MatFileError) as e:
        cm.load(str(dense))
    assert e.value.code == 2


def test_writer_rejects_bad_graphs(tmp_path):
    from graphs_util import csr_raw
    G = csr_raw(3, [[5], [], []])  # column id out of range
    with pytest.raises(graphs.MatFileError) as e:
        cm.save_mat(str(tmp_path / "x.mat"), G)
    assert e.value.code == 4
    assert not os.path.exists(tmp_path / "x.mat")
