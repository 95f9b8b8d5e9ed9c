"""Seeded benchmark graphs built directly in HBM.

Same graphs as ``graphs.py`` (pair for pair, then the same symmetric simple
CSR), built on the device: pair generation, expansion to both directions,
onesweep radix sort, dedup and row offsets.  Used for inputs too large to
build on the host in reasonable time (the 65.6M-vertex / 1.8B-pair power-law
graph takes ~43 GB of keys while it is being built).
"""
from __future__ import annotations

import ctypes
import struct

import numpy as np

from . import _native as nat
from .cc import DeviceBuffer, DeviceGraph

_KIND = {"er": 0, "rmat": 1, "grid": 2, "chung_lu": 3}


def _gpu():
    lib = nat.gpu_lib()
    if not hasattr(lib, "_devgen_sigs"):
        lib.cc_hip_gen_pairs_device.restype = ctypes.c_int
        lib.cc_hip_gen_pairs_device.argtypes = [ctypes.c_int, nat.P_u64, ctypes.c_void_p,
                                                ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                                nat.P_u64, ctypes.c_void_p]
        lib.cc_hip_csr_from_pairs_device.restype = ctypes.c_int
        lib.cc_hip_csr_from_pairs_device.argtypes = [ctypes.c_uint32, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_uint64,
                                                     ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p, ctypes.c_void_p,
                                                     ctypes.c_void_p, nat.P_u64, ctypes.c_void_p]
        lib._devgen_sigs = True
    return lib


def _host():
    lib = nat.host_lib()
    if not hasattr(lib, "_devgen_sigs"):
        lib.ccgen_permutation.restype = ctypes.c_int
        lib.ccgen_permutation.argtypes = [ctypes.c_uint32, ctypes.c_uint64, nat.P_u32]
        lib.ccgen_chung_lu_cdf.restype = ctypes.c_double
        lib.ccgen_chung_lu_cdf.argtypes = [ctypes.c_uint32, ctypes.c_double,
                                           ctypes.POINTER(ctypes.c_double)]
        lib._devgen_sigs = True
    return lib


def max_pairs_of(kind: str, params: list[int]) -> int:
    """Upper bound on the undirected pairs the generator of `kind` emits."""
    return {"er": params[1], "rmat": params[1] << params[0],
            "grid": 2 * params[0] * params[1], "chung_lu": params[1]}[kind]


def footprint(kind: str, n: int, params: list[int], permute: bool = True) -> tuple[int, int]:
    """(resident, peak) device bytes of `build(kind, n, params, ...)`: what
    the finished graph holds (row offsets, the column buffer, which is the
    pair buffer reused, and the label array) and the most the build holds at
    once (pair buffer, two key buffers of both directions, row offsets,
    permutation, Chung-Lu CDF)."""
    P = max_pairs_of(kind, params)
    resident = 8 * (n + 1) + max(1, 8 * P) + max(4, 4 * n)
    peak = (8 * P + 2 * 16 * P + 8 * (n + 1) + (4 * n if permute else 0)
            + (8 * n if kind == "chung_lu" else 0))
    return resident, max(resident, peak)


def build(kind: str, n: int, params: list[int], seed: int, permute: bool,
          cdf: np.ndarray | None = None) -> DeviceGraph:
    """Generate pairs of `kind` on the device and build the CSR there."""
    gpu = _gpu()
    max_pairs = max_pairs_of(kind, params)
    pairbuf = DeviceBuffer(8 * max_pairs)  # src | dst, later reused for col_idx
    d_src = pairbuf.ptr
    d_dst = pairbuf.ptr + 4 * max_pairs
    d_cdf = DeviceBuffer.from_array(cdf) if cdf is not None else None
    p = np.array(params, dtype=np.uint64)
    cnt = ctypes.c_uint64(0)
    nat.check(gpu.cc_hip_gen_pairs_device(_KIND[kind], nat.ptr(p, ctypes.c_uint64),
                                          d_cdf.ptr if d_cdf else None, seed, d_src, d_dst,
                                          ctypes.byref(cnt), None), "cc_hip_gen_pairs_device")
    if d_cdf:
        d_cdf.close()
    d_perm = None
    if permute:
        perm = np.empty(n, dtype=np.uint32)
        if _host().ccgen_permutation(n, seed, nat.ptr(perm, ctypes.c_uint32)) != 0:
            raise MemoryError("ccgen_permutation")
        d_perm = DeviceBuffer.from_array(perm)
        del perm
    keys = DeviceBuffer(16 * cnt.value)
    alt = DeviceBuffer(16 * cnt.value)
    row_ptr = DeviceBuffer(8 * (n + 1))
    m = ctypes.c_uint64(0)
    # col_idx is written into the pair buffer once the pairs are expanded
    nat.check(gpu.cc_hip_csr_from_pairs_device(n, d_src, d_dst, cnt.value,
                                               d_perm.ptr if d_perm else None, keys.ptr, alt.ptr,
                                               row_ptr.ptr, pairbuf.ptr, ctypes.byref(m), None),
              "cc_hip_csr_from_pairs_device")
    keys.close()
    alt.close()
    if d_perm:
        d_perm.close()
    return DeviceGraph.adopt(n, int(m.value), row_ptr, pairbuf)


def rmat_params(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19,
                c: float = 0.19) -> tuple[int, list[int]]:
    ta, tab, tabc = int(a * 65536.0), int((a + b) * 65536.0), int((a + b + c) * 65536.0)
    return 1 << scale, [scale, edge_factor, ta, tab, tabc]


def er_params(n: int, avg_degree: float) -> tuple[int, list[int]]:
    return n, [n, int(round(n * avg_degree / 2))]


def grid_params(rows: int, cols: int, keep: float = 0.7) -> tuple[int, list[int]]:
    return rows * cols, [rows, cols, int(keep * 65536.0)]


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1002, permute: bool = True) -> DeviceGraph:
    n, params = rmat_params(scale, edge_factor, a, b, c)
    return build("rmat", n, params, seed, permute)


def erdos_renyi(n: int, avg_degree: float, seed: int) -> DeviceGraph:
    n, params = er_params(n, avg_degree)
    return build("er", n, params, seed, False)


def grid(rows: int, cols: int, keep: float = 0.7, seed: int = 1004,
         permute: bool = True) -> DeviceGraph:
    n, params = grid_params(rows, cols, keep)
    return build("grid", n, params, seed, permute)


def chung_lu(n: int, pairs: int, gamma: float = 2.1, seed: int = 1003,
             permute: bool = True) -> DeviceGraph:
    cdf = np.empty(n, dtype=np.float64)
    total = _host().ccgen_chung_lu_cdf(n, gamma, cdf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    bits = struct.unpack("<Q", struct.pack("<d", total))[0]
    return build("chung_lu", n, [n, pairs, bits], seed, permute, cdf=cdf)
