"""Local connected components: Python mirror of the reference interface.

``compute_connected_components_pthreads`` has the argument meaning of the
reference's C entry point (include/cc.h:27-28, src/cc_pthreads.c:152-229) and
calls the drop-in symbol of ``libccmi355x.so`` through its C ABI.
``DeviceGraph`` keeps a CSR resident in HBM and solves it with
``cc_hip_components_device`` (no host round trip per solve).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .graphs import Graph

DEFAULT_CHUNK_SIZE = 4096  # include/cc.h:7


def compute_connected_components_pthreads(G: Graph, labels: np.ndarray | None = None,
                                          num_threads: int = 0,
                                          chunk_size: int = DEFAULT_CHUNK_SIZE) -> np.ndarray:
    """labels[v] = smallest vertex id in v's component (computed on the GPU).

    ``num_threads`` and ``chunk_size`` are accepted for interface parity and do
    not change the result.  Returns ``labels`` (allocated if None).
    """
    lib = nat.gpu_lib()
    if labels is None:
        labels = np.empty(G.n, dtype=np.uint32)
    if labels.dtype != np.uint32 or labels.size != G.n or not labels.flags["C_CONTIGUOUS"]:
        raise ValueError("labels must be a contiguous uint32 array of n entries")
    if G.n == 0:
        return labels
    rc = lib.cc_hip_components_host(nat.ptr(G.row_ptr, ctypes.c_uint64),
                                    nat.ptr(G.col_idx, ctypes.c_uint32), G.n, G.m,
                                    nat.ptr(labels, ctypes.c_uint32))
    nat.check(rc, "cc_hip_components_host")
    return labels


class DeviceBuffer:
    """A device allocation owned by Python (freed on close/GC)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        nat.check(nat.gpu_lib().cc_hip_malloc(ctypes.byref(p), max(1, self.nbytes)),
                  "cc_hip_malloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray, stream=None) -> None:
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("upload larger than buffer")
        nat.check(nat.gpu_lib().cc_hip_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes, stream),
                  "cc_hip_memcpy_h2d")

    def download(self, out: np.ndarray, stream=None) -> np.ndarray:
        if out.nbytes > self.nbytes or not out.flags["C_CONTIGUOUS"]:
            raise ValueError("bad download target")
        nat.check(nat.gpu_lib().cc_hip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes, stream),
                  "cc_hip_memcpy_d2h")
        return out

    def close(self) -> None:
        if self.ptr:
            nat.gpu_lib().cc_hip_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceGraph:
    """A CSR resident in device memory plus a device label array."""

    def __init__(self, G: Graph | None):
        if G is None:
            return
        self.n, self.m = G.n, G.m
        self.row_ptr = DeviceBuffer.from_array(G.row_ptr)
        self.col_idx = DeviceBuffer.from_array(G.col_idx) if G.m else DeviceBuffer(4)
        self.labels = DeviceBuffer(max(4, 4 * G.n))

    @classmethod
    def adopt(cls, n: int, m: int, row_ptr: "DeviceBuffer", col_idx: "DeviceBuffer") -> "DeviceGraph":
        """Wrap device arrays built elsewhere (takes ownership)."""
        self = cls(None)
        self.n, self.m = n, m
        self.row_ptr, self.col_idx = row_ptr, col_idx
        self.labels = DeviceBuffer(max(4, 4 * n))
        return self

    def to_host(self) -> Graph:
        """Copy the CSR back to host memory."""
        rp = np.empty(self.n + 1, dtype=np.uint64)
        ci = np.empty(self.m, dtype=np.uint32)
        self.row_ptr.download(rp)
        if self.m:
            self.col_idx.download(ci)
        return Graph(rp, ci)

    def solve_async(self, stream=None) -> None:
        """Enqueue a solve on `stream` (None = the library stream)."""
        rc = nat.gpu_lib().cc_hip_components_device(self.row_ptr.ptr, self.col_idx.ptr, self.n,
                                                     self.m, self.labels.ptr, stream)
        nat.check(rc, "cc_hip_components_device")

    def solve(self, stream=None) -> np.ndarray:
        self.solve_async(stream)
        out = np.empty(self.n, dtype=np.uint32)
        return self.labels.download(out, stream)

    def close(self) -> None:
        for b in (self.row_ptr, self.col_idx, self.labels):
            b.close()
