"""ORACLE — test infrastructure only.

Python bindings to ``libccoracle.so`` (this project's C restatement of the
reference path, ``cc_oracle.c``) and, when built, ``_ref/libccref.so`` (the
reference's own ``src/cc_pthreads.c`` and ``src/boundary_sort.c`` compiled
from its sources by ``oracle/Makefile``; ``_ref/libccref_mpi.so``, its
``src/exchange_plan.c``, is driven by ``tests/mpi/delta_check.c``).  The
``_ref`` builds exist only in the build container.  Only tests, ``smoke()``
and ``bench.py`` (its CPU baseline and its parity checks) may import this
package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

c_u32, c_u64, c_int = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
P_u32 = ctypes.POINTER(ctypes.c_uint32)
P_u64 = ctypes.POINTER(ctypes.c_uint64)
P_int = ctypes.POINTER(ctypes.c_int)


class _CSR(ctypes.Structure):
    _fields_ = [("n", c_u32), ("m", c_u64), ("row_ptr", P_u64), ("col_idx", P_u32)]


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libccoracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} missing; run make -C oracle")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        L.oracle_cc_min_label.argtypes = [c_u32, P_u64, P_u32, P_u32]
        L.oracle_cc_min_label.restype = None
        L.oracle_cc_lp_pthreads.argtypes = [c_u32, P_u64, P_u32, P_u32, c_int, c_int]
        L.oracle_cc_lp_pthreads.restype = c_int
        L.oracle_set_verbose.argtypes = [c_int]
        L.oracle_set_verbose.restype = None
        L.oracle_sort_u64.argtypes = [P_u64, c_u64]
        L.oracle_sort_u64.restype = c_int
        L.oracle_dedup_u64.argtypes = [P_u64, c_u64]
        L.oracle_dedup_u64.restype = c_u64
        L.oracle_ghost_extract.argtypes = [P_u64, c_u64, P_u32]
        L.oracle_ghost_extract.restype = c_u32
        L.oracle_delta_pack.argtypes = [P_u32, P_int, c_int, P_u32, P_u32, c_u32, c_u32, P_u32,
                                        P_u64, P_int, P_int]
        L.oracle_delta_pack.restype = c_u64
        L.oracle_delta_unpack.argtypes = [P_u64, P_int, P_int, P_int, c_int, P_u32, P_u32]
        L.oracle_delta_unpack.restype = None
        L.oracle_vertex_range.argtypes = [c_u32, c_int, c_int, P_u32, P_u32]
        L.oracle_vertex_range.restype = None
        L.oracle_weighted_bounds.argtypes = [P_u64, c_u32, c_int, c_u32, P_u32]
        L.oracle_weighted_bounds.restype = None
        L.oracle_owner_of_vertex.argtypes = [c_u32, c_u32, c_int, P_u32, c_int]
        L.oracle_owner_of_vertex.restype = c_int
        L.oracle_local_split.argtypes = [c_u32, P_u64, P_u32, c_u32, c_u32, c_u32, P_u64, P_u32,
                                         P_u64]
        L.oracle_local_split.restype = c_u64
        L.oracle_boundary_build.argtypes = [c_u32, P_u64, P_u32, c_u32, c_u32, c_u32, P_u32,
                                            P_u64]
        L.oracle_boundary_build.restype = c_u64
        L.oracle_merge_step.argtypes = [P_u64, c_u64, P_u32, P_u32, c_u32, P_u32, P_u32, P_u32,
                                        P_u32, c_u32, c_u32, c_u32]
        L.oracle_merge_step.restype = c_int
        _lib = L
    return _lib


def ref_lib():
    """The reference compiled from its own sources, or None if not built."""
    global _ref
    if _ref is None:
        path = os.path.join(_HERE, "_ref", "libccref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        L.compute_connected_components_pthreads.argtypes = [ctypes.POINTER(_CSR), P_u32, c_int,
                                                            c_int]
        L.compute_connected_components_pthreads.restype = None
        L.boundary_edges_sort.argtypes = [P_u64, c_u64]
        L.boundary_edges_sort.restype = None
        L.boundary_edges_dedup.argtypes = [P_u64, c_u64]
        L.boundary_edges_dedup.restype = c_u64
        _ref = L
    return _ref


# ---- convenience wrappers ------------------------------------------------

def cc_min_label(row_ptr: np.ndarray, col_idx: np.ndarray) -> np.ndarray:
    n = row_ptr.size - 1
    out = np.empty(n, dtype=np.uint32)
    lib().oracle_cc_min_label(n, _p(row_ptr, c_u64), _p(col_idx, c_u32), _p(out, c_u32))
    return out


def cc_lp_pthreads(row_ptr, col_idx, num_threads=1, chunk_size=4096) -> np.ndarray:
    n = row_ptr.size - 1
    out = np.empty(n, dtype=np.uint32)
    rc = lib().oracle_cc_lp_pthreads(n, _p(row_ptr, c_u64), _p(col_idx, c_u32), _p(out, c_u32),
                                     num_threads, chunk_size)
    if rc != 0:
        raise MemoryError("oracle_cc_lp_pthreads")
    return out


def ref_cc_pthreads(row_ptr, col_idx, num_threads=1, chunk_size=4096) -> np.ndarray:
    L = ref_lib()
    if L is None:
        raise RuntimeError("oracle/_ref/libccref.so not built")
    n = row_ptr.size - 1
    out = np.empty(n, dtype=np.uint32)
    G = _CSR(n, col_idx.size, _p(row_ptr, c_u64), _p(col_idx, c_u32))
    L.compute_connected_components_pthreads(ctypes.byref(G), _p(out, c_u32), num_threads,
                                            chunk_size)
    return out


def sort_u64(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    if lib().oracle_sort_u64(_p(k, c_u64), k.size) != 0:
        raise MemoryError("oracle_sort_u64")
    return k


def dedup_u64(keys: np.ndarray) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    cnt = lib().oracle_dedup_u64(_p(k, c_u64), k.size)
    return k[:cnt]


def ghost_extract(records: np.ndarray):
    r = np.ascontiguousarray(records, dtype=np.uint64).copy()
    ghosts = np.empty(max(1, r.size), dtype=np.uint32)
    g = lib().oracle_ghost_extract(_p(r, c_u64), r.size, _p(ghosts, c_u32))
    return r, ghosts[:g].copy()


def delta_pack(send_to_flat, seg_start, comp_label, comp_of, v_start, v_end, prev_sent):
    """Returns (pairs, sendcounts, sdispls); prev_sent is updated in place."""
    nseg = len(seg_start) - 1
    seg = np.ascontiguousarray(seg_start, dtype=np.int32)
    stf = np.ascontiguousarray(send_to_flat, dtype=np.uint32)
    cl = np.ascontiguousarray(comp_label, dtype=np.uint32)
    co = np.ascontiguousarray(comp_of, dtype=np.uint32)
    buf = np.empty(max(1, stf.size), dtype=np.uint64)
    cnt = np.zeros(max(1, nseg), dtype=np.int32)
    dis = np.zeros(max(1, nseg), dtype=np.int32)
    tot = lib().oracle_delta_pack(_p(stf, c_u32), _p(seg, c_int), nseg, _p(cl, c_u32),
                                  _p(co, c_u32), v_start, v_end, _p(prev_sent, c_u32),
                                  _p(buf, c_u64), _p(cnt, c_int), _p(dis, c_int))
    return buf[:tot].copy(), cnt[:nseg].copy(), dis[:nseg].copy()


def delta_unpack(recvbuf, recvcounts, rdispls, need_base, need_from_gidx_flat, ghost_labels):
    """oracle_delta_unpack; ghost_labels (uint32) is updated in place."""
    rb = np.ascontiguousarray(recvbuf, dtype=np.uint64)
    rc = np.ascontiguousarray(recvcounts, dtype=np.int32)
    rd = np.ascontiguousarray(rdispls, dtype=np.int32)
    nb = np.ascontiguousarray(need_base, dtype=np.int32)
    gx = np.ascontiguousarray(need_from_gidx_flat, dtype=np.uint32)
    if rc.size != rd.size or rc.size != nb.size:
        raise ValueError("recvcounts, rdispls and need_base need one entry per source")
    for k in range(rc.size):  # bounds of every read and write, before the C loop
        if rc[k] == 0:
            continue
        if rd[k] < 0 or rd[k] + rc[k] > rb.size:
            raise ValueError(f"source {k}: pairs outside recvbuf")
        pos = (rb[rd[k]:rd[k] + rc[k]] >> np.uint64(32)).astype(np.int64)
        if nb[k] < 0 or nb[k] + int(pos.max()) >= gx.size:
            raise ValueError(f"source {k}: positions outside need_from_gidx_flat")
        if int(gx[nb[k] + pos].max()) >= ghost_labels.size:
            raise ValueError(f"source {k}: ghost index outside ghost_labels")
    if rb.size == 0:
        rb = np.zeros(1, dtype=np.uint64)
    if gx.size == 0:
        gx = np.zeros(1, dtype=np.uint32)
    lib().oracle_delta_unpack(_p(rb, c_u64), _p(rc, c_int), _p(rd, c_int), _p(nb, c_int),
                              rc.size, _p(gx, c_u32), _p(ghost_labels, c_u32))


def merge_step(records, ghost_vertices, ghost_labels, comp_of, comp_label, comp_min, rep_list,
               v_start, v_end) -> int:
    """oracle_merge_step; comp_label and comp_min (uint32) are updated in
    place.  Returns 1 if any comp_label dropped."""
    rec = np.ascontiguousarray(records, dtype=np.uint64)
    gv = np.ascontiguousarray(ghost_vertices, dtype=np.uint32)
    gl = np.ascontiguousarray(ghost_labels, dtype=np.uint32)
    co = np.ascontiguousarray(comp_of, dtype=np.uint32)
    rl = np.ascontiguousarray(rep_list, dtype=np.uint32)
    z64, z32 = np.zeros(1, np.uint64), np.zeros(1, np.uint32)
    return int(lib().oracle_merge_step(
        _p(rec if rec.size else z64, c_u64), rec.size, _p(gv if gv.size else z32, c_u32),
        _p(gl if gl.size else z32, c_u32), gv.size, _p(co, c_u32), _p(comp_label, c_u32),
        _p(comp_min, c_u32), _p(rl, c_u32), rl.size, v_start, v_end))
