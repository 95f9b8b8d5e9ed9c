"""Device-built benchmark graphs equal the host-built ones exactly (same
pairs from the same seeded counters, same symmetric simple CSR), and the
solve on a device-built graph matches the oracle."""
import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import devgen
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["er", "rmat", "grid", "chung_lu", "rmat_noperm"])
def test_device_build_equals_host_build(gpu, kind):
    if kind == "er":
        d, h = devgen.erdos_renyi(50000, 6.0, seed=3), cm.erdos_renyi(50000, 6.0, seed=3)
    elif kind == "rmat":
        d, h = devgen.rmat(15, 16, seed=4), cm.rmat(15, 16, seed=4)
    elif kind == "rmat_noperm":
        d, h = devgen.rmat(12, 8, seed=5, permute=False), cm.rmat(12, 8, seed=5, permute=False)
    elif kind == "grid":
        d, h = devgen.grid(200, 300, 0.7, seed=6), cm.grid(200, 300, 0.7, seed=6)
    else:
        d, h = devgen.chung_lu(100000, 400000, 2.1, seed=7), cm.chung_lu(100000, 400000, 2.1,
                                                                          seed=7)
    g = d.to_host()
    assert g.n == h.n and g.m == h.m
    assert np.array_equal(g.row_ptr, h.row_ptr)
    assert np.array_equal(g.col_idx, h.col_idx)
    assert np.array_equal(d.solve(), oracle.cc_min_label(h.row_ptr, h.col_idx))
    d.close()


def test_device_build_edge_cases(gpu):
    # all self-loops: no edges at all
    d = devgen.grid(1, 1, 1.0, seed=1, permute=False)
    g = d.to_host()
    assert g.n == 1 and g.m == 0 and g.row_ptr.tolist() == [0, 0]
    # isolated tail vertices
    d = devgen.erdos_renyi(1000, 0.1, seed=2)
    h = cm.erdos_renyi(1000, 0.1, seed=2)
    g = d.to_host()
    assert np.array_equal(g.row_ptr, h.row_ptr) and np.array_equal(g.col_idx, h.col_idx)
