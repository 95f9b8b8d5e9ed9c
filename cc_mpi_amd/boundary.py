"""Boundary records: Python mirror of include/boundary_sort.h and the
device-resident record operations (sort, dedup, ghost extraction) and the
delta pack of the halo exchange.

A BoundaryEdge {uint32 rep; uint32 remote} array is handled as a uint64
array whose values are (remote << 32) | rep (little-endian layout; the
reference's sort key, src/boundary_sort.c:5-8).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .cc import DeviceBuffer


def _records(edges) -> np.ndarray:
    a = np.asarray(edges)
    if a.dtype != np.uint64 or not a.flags["C_CONTIGUOUS"] or not a.flags["WRITEABLE"]:
        raise ValueError("records must be a writeable contiguous uint64 array")
    return a


def pack_records(rep, remote) -> np.ndarray:
    """BoundaryEdge records from rep/remote arrays."""
    return (np.asarray(remote, dtype=np.uint64) << np.uint64(32)) | np.asarray(rep, dtype=np.uint64)


def boundary_edges_sort(edges: np.ndarray) -> None:
    """In-place ascending sort (reference src/boundary_sort.c:93-98)."""
    a = _records(edges)
    if a.size <= 1:
        return
    nat.gpu_lib().boundary_edges_sort(nat.ptr(a, ctypes.c_uint64), a.size)


def boundary_edges_dedup(edges: np.ndarray) -> int:
    """In-place adjacent-unique; returns the new count (src/boundary_sort.c:100-119)."""
    a = _records(edges)
    return int(nat.gpu_lib().boundary_edges_dedup(nat.ptr(a, ctypes.c_uint64), a.size))


def sort_device(d_keys: DeviceBuffer, d_alt: DeviceBuffer, count: int, stream=None) -> int:
    """Sort count keys resident on the device; returns 1 if the result is in d_alt."""
    which = ctypes.c_int(0)
    nat.check(nat.gpu_lib().cc_hip_sort_u64_device(d_keys.ptr, d_alt.ptr, count,
                                                   ctypes.byref(which), stream),
              "cc_hip_sort_u64_device")
    return which.value


def dedup_device(d_in: DeviceBuffer, count: int, d_out: DeviceBuffer, stream=None) -> int:
    out = ctypes.c_uint64(0)
    nat.check(nat.gpu_lib().cc_hip_dedup_u64_device(d_in.ptr, count, d_out.ptr,
                                                    ctypes.byref(out), stream),
              "cc_hip_dedup_u64_device")
    return int(out.value)


def ghost_extract_device(d_records: DeviceBuffer, count: int, d_out: DeviceBuffer,
                         d_ghosts: DeviceBuffer, stream=None) -> int:
    g = ctypes.c_uint64(0)
    nat.check(nat.gpu_lib().cc_hip_ghost_extract_device(d_records.ptr, count, d_out.ptr,
                                                        d_ghosts.ptr, ctypes.byref(g), stream),
              "cc_hip_ghost_extract_device")
    return int(g.value)


class PinnedBuffer:
    """Pinned, device-visible host memory viewed as a numpy array."""

    def __init__(self, count: int, dtype=np.uint64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = ctypes.c_void_p()
        nat.check(nat.gpu_lib().cc_hip_host_alloc(ctypes.byref(p),
                                                  max(1, self.count * self.dtype.itemsize)),
                  "cc_hip_host_alloc")
        self.ptr = p.value
        buf = (ctypes.c_char * max(1, self.count * self.dtype.itemsize)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.count)

    def close(self):
        if self.ptr:
            self.array = None
            nat.gpu_lib().cc_hip_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def delta_pack_device(d_send_to_flat: DeviceBuffer, total_send_to: int, seg_start,
                      d_comp_label: DeviceBuffer, d_comp_of: DeviceBuffer, v_start: int,
                      v_end: int, d_prev_sent: DeviceBuffer, sendbuf: PinnedBuffer, stream=None,
                      d_rep_of_send: DeviceBuffer | None = None):
    """Delta pack (src/exchange_plan.c:325-357) on device-resident arrays.
    With d_rep_of_send (delta_prepare_device), the send entries' reps are
    read from it instead of gathered from d_comp_of.

    Returns (pairs view into sendbuf, sendcounts, sdispls)."""
    seg = np.ascontiguousarray(seg_start, dtype=np.int32)
    nseg = seg.size - 1
    cnt = np.zeros(max(1, nseg), dtype=np.int32)
    dis = np.zeros(max(1, nseg), dtype=np.int32)
    tot = ctypes.c_uint64(0)
    lib = nat.gpu_lib()
    if d_rep_of_send is not None:
        nat.check(lib.cc_hip_delta_pack_reps_device(
            d_send_to_flat.ptr, d_rep_of_send.ptr, total_send_to, nat.ptr(seg, ctypes.c_int), nseg,
            d_comp_label.ptr, d_prev_sent.ptr, sendbuf.ptr, nat.ptr(cnt, ctypes.c_int),
            nat.ptr(dis, ctypes.c_int), ctypes.byref(tot), stream), "cc_hip_delta_pack_reps_device")
    else:
        nat.check(lib.cc_hip_delta_pack_device(
            d_send_to_flat.ptr, total_send_to, nat.ptr(seg, ctypes.c_int), nseg, d_comp_label.ptr,
            d_comp_of.ptr, v_start, v_end, d_prev_sent.ptr, sendbuf.ptr,
            nat.ptr(cnt, ctypes.c_int), nat.ptr(dis, ctypes.c_int), ctypes.byref(tot), stream),
            "cc_hip_delta_pack_device")
    return sendbuf.array[:int(tot.value)], cnt[:nseg], dis[:nseg]


def delta_prepare_device(d_send_to_flat: DeviceBuffer, total_send_to: int,
                         d_comp_of: DeviceBuffer, v_start: int, v_end: int,
                         d_rep_of_send: DeviceBuffer, stream=None) -> None:
    """The send entries' reps gathered once (cc_hip_delta_prepare_device)."""
    nat.check(nat.gpu_lib().cc_hip_delta_prepare_device(
        d_send_to_flat.ptr, total_send_to, d_comp_of.ptr, v_start, v_end, d_rep_of_send.ptr,
        stream), "cc_hip_delta_prepare_device")
