"""ctypes bindings to the in-tree native libraries.

``libccmi355x.so`` is the HIP library (kernels + C ABI, gfx950 only) and
``libcchost.so`` the host C library (generators, CSR builder).  Both are built
in-tree by ``__graft_entry__.build()`` (``make -C cc_mpi_amd/csrc``).  Nothing
here falls back to a CPU implementation: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(_HERE, "lib")

c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_int = ctypes.c_int
c_void_p = ctypes.c_void_p
c_size_t = ctypes.c_size_t
c_double = ctypes.c_double
P_u32 = ctypes.POINTER(ctypes.c_uint32)
P_u64 = ctypes.POINTER(ctypes.c_uint64)
P_int = ctypes.POINTER(ctypes.c_int)


class NativeLibraryMissing(RuntimeError):
    pass


class CSRGraphC(ctypes.Structure):
    """ABI twin of CSRGraph (include/graph.h; reference include/graph.h:8-14)."""

    _fields_ = [
        ("n", c_u32),
        ("m", c_u64),
        ("row_ptr", P_u64),
        ("col_idx", P_u32),
    ]


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise NativeLibraryMissing(
            f"{path} is not built; run __graft_entry__.build() (make -C cc_mpi_amd/csrc)"
        )
    return ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)


_gpu = None
_host = None


def _sig(lib, name, restype, argtypes):
    fn = getattr(lib, name)
    fn.restype = restype
    fn.argtypes = argtypes
    return fn


def gpu_lib() -> ctypes.CDLL:
    """The HIP library with argument types declared (loads once)."""
    global _gpu
    if _gpu is not None:
        return _gpu
    lib = _load("libccmi355x.so")
    _sig(lib, "cc_hip_init", c_int, [c_int])
    _sig(lib, "cc_hip_device_count", c_int, [])
    _sig(lib, "cc_hip_last_error", ctypes.c_char_p, [])
    _sig(lib, "cc_hip_build_info", ctypes.c_char_p, [])
    _sig(lib, "cc_hip_malloc", c_int, [ctypes.POINTER(c_void_p), c_size_t])
    _sig(lib, "cc_hip_free", c_int, [c_void_p])
    _sig(lib, "cc_hip_host_alloc", c_int, [ctypes.POINTER(c_void_p), c_size_t])
    _sig(lib, "cc_hip_host_free", c_int, [c_void_p])
    for nm in ("cc_hip_memcpy_h2d", "cc_hip_memcpy_d2h", "cc_hip_memcpy_d2d"):
        _sig(lib, nm, c_int, [c_void_p, c_void_p, c_size_t, c_void_p])
    _sig(lib, "cc_hip_stream_create", c_int, [ctypes.POINTER(c_void_p)])
    _sig(lib, "cc_hip_stream_destroy", c_int, [c_void_p])
    _sig(lib, "cc_hip_stream_sync", c_int, [c_void_p])
    _sig(lib, "cc_hip_device_sync", c_int, [])
    _sig(lib, "cc_hip_prof_enable", c_int, [c_int])
    _sig(lib, "cc_hip_prof_reset", c_int, [])
    _sig(lib, "cc_hip_prof_query", c_int, [ctypes.c_char_p, P_u64, ctypes.POINTER(c_double)])
    _sig(lib, "cc_hip_prof_names", ctypes.c_char_p, [])
    _sig(lib, "cc_hip_components_device", c_int,
         [c_void_p, c_void_p, c_u32, c_u64, c_void_p, c_void_p])
    _sig(lib, "cc_hip_components_host", c_int, [P_u64, P_u32, c_u32, c_u64, P_u32])
    _sig(lib, "cc_hip_components_tune", c_int, [c_int, c_int, c_int])
    _sig(lib, "cc_hip_components_set", c_int, [ctypes.c_char_p, c_int])
    _sig(lib, "cc_hip_components_get", c_int, [ctypes.c_char_p])
    _sig(lib, "cc_hip_components_plan", c_int, [c_u32, c_u64, P_int, P_int, P_int, P_int])
    _sig(lib, "cc_hip_components_plan_get", c_int, [c_u32, c_u64, ctypes.c_char_p])
    _sig(lib, "cc_hip_components_last_pass_ran", c_int, [c_void_p, P_int])
    _sig(lib, "cc_hip_components_tiles_ran", c_int, [c_void_p, P_int])
    _sig(lib, "compute_connected_components_pthreads", None,
         [ctypes.POINTER(CSRGraphC), P_u32, c_int, c_int])
    _sig(lib, "cc_hip_sort_u64_device", c_int,
         [c_void_p, c_void_p, c_u64, ctypes.POINTER(c_int), c_void_p])
    _sig(lib, "cc_hip_boundary_build_device", c_int,
         [c_void_p, c_void_p, c_u64, c_void_p, c_u32, P_u64, P_int, c_void_p])
    _sig(lib, "cc_hip_sort_set", c_int, [ctypes.c_char_p, c_int])
    _sig(lib, "cc_hip_sort_get", c_int, [ctypes.c_char_p])
    _sig(lib, "cc_hip_dedup_u64_device", c_int, [c_void_p, c_u64, c_void_p, P_u64, c_void_p])
    _sig(lib, "cc_hip_ghost_extract_device", c_int,
         [c_void_p, c_u64, c_void_p, c_void_p, P_u64, c_void_p])
    _sig(lib, "cc_hip_delta_pack_device", c_int,
         [c_void_p, c_u64, P_int, c_int, c_void_p, c_void_p, c_u32, c_u32, c_void_p, c_void_p,
          P_int, P_int, P_u64, c_void_p])
    _sig(lib, "boundary_edges_sort", None, [P_u64, c_u64])
    _sig(lib, "boundary_edges_dedup", c_u64, [P_u64, c_u64])
    # per-rank driver building blocks
    V = c_void_p
    _sig(lib, "cc_hip_split_count_device", c_int,
         [V, V, c_u32, c_u32, c_u32, c_u32, P_u64, P_u64, V])
    _sig(lib, "cc_hip_split_emit_device", c_int,
         [V, V, c_u32, c_u32, c_u32, c_u32, V, V, V, V, V])
    _sig(lib, "cc_hip_boundary_set_rep_device", c_int, [V, c_u64, V, V])
    _sig(lib, "cc_hip_iota_u32_device", c_int, [V, c_u32, c_u32, V])
    _sig(lib, "cc_hip_merge_step_device", c_int,
         [V, c_u64, V, V, c_u32, V, V, V, c_u32, c_u32, c_u32, P_int, V])
    _sig(lib, "cc_hip_merge_step_reps_device", c_int,
         [V, c_u64, V, V, c_u32, V, V, V, c_u32, V, c_u32, c_u32, c_u32, P_int, V])
    _sig(lib, "cc_hip_merge_prepare_device", c_int,
         [V, c_u64, c_u32, V, c_u32, V, V, P_u64, P_u32, V])
    _sig(lib, "cc_hip_merge_step_prepared_device", c_int,
         [V, c_u32, V, c_u64, V, V, c_u32, V, V, V, V, c_u32, V, c_u32, c_u32, c_u32, c_int,
          P_int, V])
    _sig(lib, "cc_hip_local_reps_device", c_int, [V, c_u32, V, P_u32, V])
    _sig(lib, "cc_hip_rank_labels_device", c_int, [V, V, c_u32, c_u32, V, P_u64, V])
    _sig(lib, "cc_hip_owner_keys_device", c_int, [V, c_u64, V, c_int, V, V])
    _sig(lib, "cc_hip_segments_device", c_int, [V, c_u64, c_int, c_u32, V, V, V])
    _sig(lib, "cc_hip_send_lists_device", c_int, [V, c_u64, c_u32, V, c_int, c_u32, P_int, V, V])
    _sig(lib, "cc_hip_gather_u32_device", c_int, [V, V, c_u64, V, V])
    _sig(lib, "cc_hip_exclusive_scan_u32_device", c_int, [V, c_u64, V, V])
    _sig(lib, "cc_hip_delta_unpack_device", c_int, [V, c_u64, P_int, P_int, c_int, V, V, V])
    _sig(lib, "cc_hip_delta_prepare_device", c_int, [V, c_u64, V, c_u32, c_u32, V, V])
    _sig(lib, "cc_hip_delta_pack_reps_device", c_int,
         [V, V, c_u64, P_int, c_int, V, V, V, P_int, P_int, P_u64, V])
    _sig(lib, "cc_hip_components_device_noalloc", c_int, [V, V, c_u32, c_u64, V, V])
    _gpu = lib
    return lib


def host_lib() -> ctypes.CDLL:
    global _host
    if _host is not None:
        return _host
    lib = _load("libcchost.so")
    _sig(lib, "free_csr", None, [ctypes.POINTER(CSRGraphC)])
    _sig(lib, "ccgen_rand_at", c_u64, [c_u64, c_u64])
    _sig(lib, "ccgen_csr_from_pairs", c_int,
         [c_u32, P_u32, P_u32, c_u64, P_u32, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "ccgen_erdos_renyi", c_int, [c_u32, c_u64, c_u64, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "ccgen_rmat", c_int,
         [c_int, c_u32, c_double, c_double, c_double, c_u64, c_int, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "ccgen_grid", c_int,
         [c_u32, c_u32, c_double, c_u64, c_int, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "ccgen_chung_lu", c_int,
         [c_u32, c_u64, c_double, c_u64, c_int, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "load_csr_from_mat", c_int, [ctypes.c_char_p, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "load_csr_from_file", c_int,
         [ctypes.c_char_p, c_int, c_int, ctypes.POINTER(CSRGraphC)])
    _sig(lib, "cc_mat_write_sparse", c_int,
         [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(CSRGraphC), c_int])
    _host = lib
    return lib


def ptr(a: np.ndarray, ctype=None):
    """ctypes pointer to a contiguous numpy array's data."""
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    if ctype is None:
        return ctypes.c_void_p(a.ctypes.data)
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def check(rc: int, what: str) -> None:
    """Raise with the library's message if a C-ABI call failed."""
    if rc != 0:
        msg = gpu_lib().cc_hip_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (rc={rc}): {msg}")
