"""The allocation-free solve, cc_hip_components_device_noalloc, through the
C ABI: its labels equal cc_hip_components_device's (and the oracle's) on
calls in a row, each followed by a smaller graph, on the library stream and
on a caller's stream; the device memory in use is the same before and after
its calls; a graph larger than its scratch array is refused."""
import ctypes

import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import _native as nat
from cc_mpi_amd import devgen
import oracle

pytestmark = pytest.mark.gpu

MAX_VERTICES = 1 << 22


def noalloc(lib):
    f = lib.cc_hip_components_device_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64,
                  ctypes.c_void_p, ctypes.c_void_p]
    return f


def free_device_bytes():
    hip = ctypes.CDLL("libamdhip64.so")  # the runtime the library runs on
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("use_stream", [False, True])
def test_noalloc_equals_device_solve(gpu, use_stream):
    lib = nat.gpu_lib()
    fn = noalloc(lib)
    stream = None
    if use_stream:
        s = ctypes.c_void_p()
        nat.check(lib.cc_hip_stream_create(ctypes.byref(s)), "stream")
        stream = s.value
    # the largest graph the scratch holds, then smaller ones
    # (the lattice numbered row by row takes the vertex-tile plan, whose
    # workspace is the larger one)
    graphs = [devgen.rmat(22, 16, seed=1002), devgen.grid(300, 200, 0.7, seed=5),
              devgen.grid(2048, 2048, 0.7, seed=6, permute=False),
              devgen.erdos_renyi(70_000, 3.0, seed=8)]
    assert graphs[0].n == MAX_VERTICES and graphs[2].n == MAX_VERTICES
    outs = [cm.DeviceBuffer(4 * G.n) for G in graphs]
    wants = [G.solve() for G in graphs]  # cc_hip_components_device, library stream
    for G, want in zip(graphs[1:], wants[1:]):
        h = G.to_host()
        assert np.array_equal(want, oracle.cc_min_label(h.row_ptr, h.col_idx))
    nat.check(lib.cc_hip_device_sync(), "sync")
    for rnd in range(2):
        for G, out, want in zip(graphs, outs, wants):
            print(f"noalloc solve n={G.n} m={G.m} stream={use_stream} pass {rnd}", flush=True)
            nat.check(lib.cc_hip_memcpy_d2d(out.ptr, G.row_ptr.ptr, 4 * G.n, None), "poison")
            nat.check(lib.cc_hip_device_sync(), "sync")
            before = free_device_bytes()
            rc = fn(G.row_ptr.ptr, G.col_idx.ptr, G.n, G.m, out.ptr, stream)
            assert rc == 0, lib.cc_hip_last_error().decode()
            nat.check(lib.cc_hip_stream_sync(stream), "sync")
            assert free_device_bytes() == before
            got = np.empty(G.n, dtype=np.uint32)
            out.download(got)
            assert np.array_equal(got, want), f"n={G.n} pass {rnd}"
    for b in outs:
        b.close()
    for G in graphs:
        G.close()
    if stream:
        nat.check(lib.cc_hip_stream_destroy(stream), "stream destroy")


def test_noalloc_refuses_graphs_beyond_its_scratch(gpu):
    lib = nat.gpu_lib()
    fn = noalloc(lib)
    buf = cm.DeviceBuffer(64)
    # refused before anything touches the pointers
    rc = fn(buf.ptr, buf.ptr, MAX_VERTICES + 1, 10, buf.ptr, None)
    assert rc == 4
    assert b"exceeds" in lib.cc_hip_last_error()
    assert fn(None, None, 5, 0, None, None) == 4  # null pointers
    assert fn(None, None, 0, 0, None, None) == 0  # nothing to do
    buf.close()


def test_noalloc_without_edges(gpu):
    lib = nat.gpu_lib()
    fn = noalloc(lib)
    rp = cm.DeviceBuffer.from_array(np.zeros(1001, dtype=np.uint64))
    out = cm.DeviceBuffer(4 * 1000)
    assert fn(rp.ptr, None, 1000, 0, out.ptr, None) == 0
    got = np.empty(1000, dtype=np.uint32)
    nat.check(lib.cc_hip_device_sync(), "sync")
    out.download(got)
    assert np.array_equal(got, np.arange(1000, dtype=np.uint32))
    rp.close()
    out.close()
