"""MI355X-native per-rank connected-components path (CC_mpi hot path).

Native libraries (built in-tree by ``__graft_entry__.build()``):
  * ``lib/libccmi355x.so`` — HIP kernels for gfx950 behind a C ABI
    (include/cc.h, include/boundary_sort.h, include/cc_hip.h);
  * ``lib/libcchost.so`` — host C: seeded generators, the CSR builder and
    the level 5 MAT-file reader/writer (include/graph.h).
"""
from .graphs import Graph, erdos_renyi, rmat, grid, chung_lu, from_pairs, load, save_mat  # noqa: F401
from .cc import (  # noqa: F401
    compute_connected_components_pthreads,
    DeviceBuffer,
    DeviceGraph,
    DEFAULT_CHUNK_SIZE,
)

__all__ = [
    "Graph", "erdos_renyi", "rmat", "grid", "chung_lu", "from_pairs", "load", "save_mat",
    "compute_connected_components_pthreads", "DeviceBuffer", "DeviceGraph",
    "DEFAULT_CHUNK_SIZE",
]
