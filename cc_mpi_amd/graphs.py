"""Host-side graphs: the CSRGraph mirror and the seeded generators.

``Graph`` mirrors the reference's ``CSRGraph`` (include/graph.h:8-14): ``n``
vertices, ``m = row_ptr[n]`` directed entries, ``row_ptr`` uint64[n+1],
``col_idx`` uint32[m].  Generators call ``libcchost.so`` and return symmetric
simple graphs with sorted rows.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as nat


@dataclass
class Graph:
    row_ptr: np.ndarray  # uint64[n+1]
    col_idx: np.ndarray  # uint32[m]
    name: str = ""

    def __post_init__(self):
        self.row_ptr = np.ascontiguousarray(self.row_ptr, dtype=np.uint64)
        self.col_idx = np.ascontiguousarray(self.col_idx, dtype=np.uint32)
        if self.row_ptr.ndim != 1 or self.row_ptr.size < 1:
            raise ValueError("row_ptr must be 1-D with n+1 entries")
        if int(self.row_ptr[-1]) != self.col_idx.size:
            raise ValueError(
                f"row_ptr[n]={int(self.row_ptr[-1])} != len(col_idx)={self.col_idx.size}")

    @property
    def n(self) -> int:
        return self.row_ptr.size - 1

    @property
    def m(self) -> int:
        return self.col_idx.size

    def as_c(self) -> nat.CSRGraphC:
        """A CSRGraph view of the arrays (valid while this Graph lives)."""
        return nat.CSRGraphC(self.n, self.m, nat.ptr(self.row_ptr, ctypes.c_uint64),
                             nat.ptr(self.col_idx, ctypes.c_uint32))


def _take(g: nat.CSRGraphC, name: str) -> Graph:
    lib = nat.host_lib()
    try:
        n, m = int(g.n), int(g.m)
        rp = np.ctypeslib.as_array(g.row_ptr, shape=(n + 1,)).copy()
        ci = (np.ctypeslib.as_array(g.col_idx, shape=(m,)).copy() if m
              else np.zeros(0, dtype=np.uint32))
    finally:
        lib.free_csr(ctypes.byref(g))
    return Graph(rp, ci, name)


def _gen(fn, *args, name: str) -> Graph:
    g = nat.CSRGraphC()
    rc = fn(*args, ctypes.byref(g))
    if rc != 0:
        raise RuntimeError(f"generator {name} failed (rc={rc})")
    return _take(g, name)


def from_pairs(n: int, src, dst, perm=None, name: str = "pairs") -> Graph:
    """Symmetric simple CSR from undirected pairs (self-loops/duplicates/out-of-range dropped)."""
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    if src.shape != dst.shape:
        raise ValueError("src and dst must have the same length")
    pp = None
    if perm is not None:
        perm = np.ascontiguousarray(perm, dtype=np.uint32)
        if perm.size != n:
            raise ValueError("perm must have n entries")
        pp = nat.ptr(perm, ctypes.c_uint32)
    return _gen(nat.host_lib().ccgen_csr_from_pairs, n, nat.ptr(src, ctypes.c_uint32),
                nat.ptr(dst, ctypes.c_uint32), src.size, pp, name=name)


def erdos_renyi(n: int, avg_degree: float, seed: int) -> Graph:
    """G(n, M) with M = n*avg_degree/2 random pairs."""
    pairs = int(round(n * avg_degree / 2))
    return _gen(nat.host_lib().ccgen_erdos_renyi, n, pairs, seed, name=f"er-{n}-d{avg_degree:g}")


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 1002, permute: bool = True) -> Graph:
    """R-MAT with 2^scale vertices and edge_factor*2^scale generated edges."""
    return _gen(nat.host_lib().ccgen_rmat, scale, edge_factor, a, b, c, seed, int(permute),
                name=f"rmat-s{scale}-ef{edge_factor}")


def grid(rows: int, cols: int, keep: float = 0.7, seed: int = 1004, permute: bool = True) -> Graph:
    """rows x cols 4-neighbour lattice with each edge kept with probability `keep`."""
    return _gen(nat.host_lib().ccgen_grid, rows, cols, keep, seed, int(permute),
                name=f"grid-{rows}x{cols}-keep{keep:g}")


def chung_lu(n: int, pairs: int, gamma: float = 2.1, seed: int = 1003,
             permute: bool = True) -> Graph:
    """Chung-Lu power-law graph with `pairs` sampled endpoint pairs."""
    return _gen(nat.host_lib().ccgen_chung_lu, n, pairs, gamma, seed, int(permute),
                name=f"chunglu-{n}-g{gamma:g}")


MAT_COMPRESS = 1   # CC_MAT_COMPRESS
MAT_PROBLEM = 2    # CC_MAT_PROBLEM
MAT_BIG_ENDIAN = 4  # CC_MAT_BIG_ENDIAN


class MatFileError(RuntimeError):
    """load_csr_from_file / load_csr_from_mat returned nonzero (``code``)."""

    def __init__(self, path: str, code: int):
        super().__init__(f"loading {path} failed (rc={code})")
        self.code = code


def load(path: str) -> Graph:
    """Read a graph the way the reference's loader does (load_csr_from_file,
    src/graph.c:14-128): ``.mat`` files only; rows of the stored sparse
    matrix become vertices."""
    g = nat.CSRGraphC()
    rc = nat.host_lib().load_csr_from_file(str(path).encode(), 1, 1, ctypes.byref(g))
    if rc != 0:
        raise MatFileError(str(path), rc)
    return _take(g, str(path))


def save_mat(path: str, g: Graph, name: str = "A", flags: int = 0) -> None:
    """Write ``g`` as a sparse logical matrix to a level 5 MAT file
    (cc_mat_write_sparse); ``flags`` is a mix of MAT_COMPRESS, MAT_PROBLEM
    and MAT_BIG_ENDIAN."""
    rc = nat.host_lib().cc_mat_write_sparse(str(path).encode(), name.encode(),
                                            ctypes.byref(g.as_c()), int(flags))
    if rc != 0:
        raise MatFileError(str(path), rc)
