"""Parity of the HIP local-CC path with the oracle (bit-exact labels).

Runs through the C ABI: the drop-in compute_connected_components_pthreads
(host buffers) and cc_hip_components_device (CSR resident in HBM).
"""
import ctypes

import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import _native as nat
import oracle
from graphs_util import (banded, complete, csr_raw, disjoint_cliques, lattice3d, one_directional,
                         path, star)

pytestmark = pytest.mark.gpu


def check_labels(G, got):
    want = oracle.cc_min_label(G.row_ptr, G.col_idx)
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, f"{G.name}: {bad.size}/{G.n} labels differ, first at {bad[:5]}"
    return want


def dropin(G):
    lab = np.full(G.n, 0xDEADBEEF, dtype=np.uint32)
    Gc = G.as_c()
    nat.gpu_lib().compute_connected_components_pthreads(ctypes.byref(Gc),
                                                        nat.ptr(lab, ctypes.c_uint32), 8, 8192)
    return lab


@pytest.mark.parametrize("make", [
    lambda: path(2), lambda: path(1000), lambda: path(5000, perm_seed=1),
    lambda: star(3000, centre=2999), lambda: star(3000, centre=0), lambda: complete(64),
    lambda: disjoint_cliques(300, 7, seed=3),
])
def test_small_structured(gpu, make):
    G = make()
    check_labels(G, dropin(G))


def test_edge_cases(gpu):
    # single vertex, no edges, isolated vertices
    G = csr_raw(1, [[]])
    assert dropin(G).tolist() == [0]
    G = csr_raw(5, [[], [], [], [], []])
    assert dropin(G).tolist() == [0, 1, 2, 3, 4]
    # non-simple but symmetric CSR: duplicates, self-loops, unsorted rows;
    # entries >= n are ignored
    G = csr_raw(6, [[1, 1, 0], [0, 0, 1], [77], [4], [3, 3, 4], [5]])
    assert dropin(G).tolist() == [0, 0, 2, 3, 3, 5]
    G = csr_raw(4, [[3, 2], [2], [1, 0], [0]])
    assert dropin(G).tolist() == [0, 0, 0, 0]


@pytest.mark.parametrize("knobs", [dict(tiles=1), dict(tiles=1, tile_shift=6, tile_pairs=1),
                                   dict(tiles=1, tile_shift=6, symmetric=0)],
                         ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()))
def test_edge_cases_vertex_tiles(gpu, knobs):
    """The hand-built CSRs above with the vertex-tile plan forced: one tile
    or several of 64 ids, with pair segments of one entry."""
    lib = nat.gpu_lib()
    try:
        for k, v in knobs.items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)
        for G, want in edge_case_graphs():
            assert dropin(G).tolist() == want
    finally:
        for k, v in dict(tiles=-1, tile_shift=0, tile_pairs=0, symmetric=1).items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)


@pytest.mark.parametrize("shift,split", [(1, 0), (2, 0), (8, 0), (1, 1), (8, 1)])
def test_edge_cases_low_first_hook(gpu, shift, split):
    """The same CSRs through the low-first hook (the one-hop skipping hook
    from the CSR with the low-id map): graphs below 64 vertices run one pass,
    the 210-vertex one hooks ids below 64 or 128 first, whose map then
    serves the rest (entries past n, self-loops and unsorted rows included),
    and a 70-vertex path whose ids above 64 all link through the map."""
    lib = nat.gpu_lib()
    knobs = dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=shift,
                 split_walks=split)
    n = 70
    path = [[] for _ in range(n)]
    for a in range(n - 1):
        path[a].append(a + 1)
        path[a + 1].append(a)
    cases = edge_case_graphs() + [(csr_raw(n, path), [0] * n)]
    try:
        for k, v in knobs.items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)
        for G, want in cases:
            assert dropin(G).tolist() == want
    finally:
        for k, v in dict(csr_hook=-1, rem=-1, halve=-1, hook_ilp=0, skip_linked=-1,
                         low_shift=-1, split_walks=-1).items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)


def edge_case_graphs():
    cases = [
        (csr_raw(1, [[]]), [0]),
        (csr_raw(5, [[], [], [], [], []]), [0, 1, 2, 3, 4]),
        (csr_raw(6, [[1, 1, 0], [0, 0, 1], [77], [4], [3, 3, 4], [5]]), [0, 0, 2, 3, 3, 5]),
        (csr_raw(4, [[3, 2], [2], [1, 0], [0]]), [0, 0, 0, 0]),
    ]
    # rows across tiles of 64: a path 0-70-130-200 and a pair 65-199,
    # duplicates, self-loops, entries past n and unsorted rows
    n = 210
    adj = [[] for _ in range(n)]
    for a, b in ((0, 70), (70, 130), (130, 200), (65, 199), (5, 6), (6, 5), (63, 64)):
        adj[a].append(b)
        adj[b].append(a)
    adj[9] += [9, 999, 9]
    adj[130] = adj[130][::-1]
    want = oracle.cc_min_label(csr_raw(n, adj).row_ptr, csr_raw(n, adj).col_idx).tolist()
    cases.append((csr_raw(n, adj), want))
    return cases


def test_empty_graph(gpu):
    G = cm.Graph(np.zeros(1, dtype=np.uint64), np.zeros(0, dtype=np.uint32))
    assert cm.compute_connected_components_pthreads(G).size == 0


def test_vertex_limit_rejected(gpu):
    """More than 0xFF000000 vertices is refused on the host (error 4) before
    any device memory is touched; null pointers likewise."""
    lib = nat.gpu_lib()
    dummy = ctypes.c_void_p(0x1000)
    assert lib.cc_hip_components_device(dummy, dummy, 0xFF000001, 1, dummy, None) == 4
    assert b"limit" in lib.cc_hip_last_error()
    assert lib.cc_hip_components_device(None, dummy, 10, 1, dummy, None) == 4


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_er_64k(gpu, seed):
    # Erdos-Renyi 65,536 vertices, average degree 8 (the reference's CPU case)
    G = cm.erdos_renyi(65536, 8.0, seed=1000 + seed)
    check_labels(G, dropin(G))


@pytest.mark.parametrize("avg", [0.5, 1.0, 2.0])
def test_er_near_threshold(gpu, avg):
    # many mid-sized components around the giant-component threshold
    G = cm.erdos_renyi(200000, avg, seed=17)
    check_labels(G, dropin(G))


@pytest.mark.parametrize("scale", [10, 14, 18])
def test_rmat(gpu, scale):
    G = cm.rmat(scale, 16, seed=1002)
    check_labels(G, dropin(G))


@pytest.mark.parametrize("keep", [0.3, 0.5, 0.7, 1.0])
def test_grid(gpu, keep):
    G = cm.grid(300, 400, keep, seed=1004)
    check_labels(G, dropin(G))


def test_chung_lu(gpu):
    G = cm.chung_lu(1 << 18, 3 << 19, 2.1, seed=1003)
    check_labels(G, dropin(G))


@pytest.mark.parametrize("kind,rem,splice,halve,skip,tiles,ran", [
    ("er", 1, 1, 0, 0, 2, 0), ("er8", 1, 1, 0, 2, 2, 0), ("grid", 1, 1, 0, 0, 2, 0),
    ("grid_rows", 1, 1, 0, 0, 2, 1), ("rmat24", 0, 0, 1, 2, 0, 0)])
def test_default_plan_at_scale(gpu, kind, rem, splice, halve, skip, tiles, ran):
    """Graphs of 2^24 vertices: sparse ones take Rem's unions with splicing,
    dense ones the lockstep hook from the CSR with path halving, skipping,
    the low-first map and queued walks, by default, and sparse
    ones of degree 6 and more skip neighbour 1 of linked vertices; on sparse
    ones a probe on the device picks the vertex-tile plan where ids are
    local (the lattice numbered row by row) and not elsewhere; labels,
    component count and size histogram equal the oracle's."""
    from cc_mpi_amd import devgen
    lib = nat.gpu_lib()
    D = (devgen.erdos_renyi(1 << 24, 3.0, seed=41) if kind == "er"
         else devgen.erdos_renyi(1 << 24, 8.0, seed=44) if kind == "er8"
         else devgen.grid(4096, 4096, 0.7, seed=42) if kind == "grid"
         else devgen.grid(4096, 4096, 0.7, seed=45, permute=False) if kind == "grid_rows"
         else devgen.rmat(24, 16, seed=43))
    plan = {k: lib.cc_hip_components_plan_get(D.n, D.m, k.encode())
            for k in ("fused", "rem", "rem_splice", "halve", "skip_linked", "tiles",
                      "csr_hook", "low_shift", "split_walks")}
    dense = kind == "rmat24"
    assert plan == dict(fused=1, rem=rem, rem_splice=splice, halve=halve, skip_linked=skip,
                        tiles=tiles, csr_hook=int(dense), low_shift=4 if dense else 0,
                        split_walks=int(dense))
    got = D.solve()
    took = ctypes.c_int(-1)
    nat.check(lib.cc_hip_components_tiles_ran(None, ctypes.byref(took)), "tiles ran")
    assert took.value == ran
    G = D.to_host()
    D.close()
    want = check_labels(G, got)
    roots = np.flatnonzero(want == np.arange(G.n))
    assert np.count_nonzero(got == np.arange(G.n)) == roots.size
    assert np.array_equal(np.bincount(got, minlength=G.n)[roots],
                          np.bincount(want, minlength=G.n)[roots])


@pytest.mark.parametrize("kind", ["rmat24_unpermuted", "dense_blocks"])
def test_low_first_plan_at_scale(gpu, kind):
    """Dense graphs of 2^24 vertices on the default plan (the low-first hook
    from the CSR with queued walks) where its map has little to work with:
    R-MAT 24 with ids not permuted, so the low ids are the hubs; and 64
    disjoint random blocks of 20 entries per vertex with permuted ids, so
    the giant root of the low ids holds about 1/64 of the vertices and most
    links find neighbour 0 unmarked.  Labels, component count and sizes
    equal the oracle's."""
    from cc_mpi_amd import devgen
    lib = nat.gpu_lib()
    if kind == "rmat24_unpermuted":
        D = devgen.rmat(24, 16, seed=47, permute=False)
        G = D.to_host()
    else:
        n, blocks = 1 << 24, 64
        size = n // blocks
        rng = np.random.default_rng(48)
        pairs = 10 * n
        base = rng.integers(0, blocks, pairs, dtype=np.uint32) * np.uint32(size)
        src = base + rng.integers(0, size, pairs, dtype=np.uint32)
        dst = base + rng.integers(0, size, pairs, dtype=np.uint32)
        G = cm.from_pairs(n, src, dst, perm=rng.permutation(n).astype(np.uint32))
        del base, src, dst
        D = cm.DeviceGraph(G)
    assert G.m >= 16 * G.n and G.m < 1 << 31
    plan = {k: lib.cc_hip_components_plan_get(G.n, G.m, k.encode())
            for k in ("csr_hook", "skip_linked", "low_shift", "split_walks")}
    assert plan == dict(csr_hook=1, skip_linked=2, low_shift=4, split_walks=1)
    got = D.solve()
    D.close()
    want = check_labels(G, got)
    roots = np.flatnonzero(want == np.arange(G.n))
    if kind == "dense_blocks":
        assert roots.size >= blocks
    assert np.array_equal(np.bincount(got, minlength=G.n)[roots],
                          np.bincount(want, minlength=G.n)[roots])


def test_last_pass_gate(gpu):
    """The last pointer-jumping pass is skipped when the frontier sweep links
    no root; labels are exact either way, and both cases occur here."""
    lib = nat.gpu_lib()
    seen = set()
    for G in (disjoint_cliques(200, 5, seed=34), cm.grid(120, 150, 0.6, seed=33),
              cm.rmat(14, 16, seed=31), cm.erdos_renyi(50000, 1.2, seed=32),
              cm.chung_lu(1 << 16, 1 << 19, 2.1, seed=35), path(3000, perm_seed=2)):
        D = cm.DeviceGraph(G)
        got = D.solve()
        ran = ctypes.c_int(-1)
        nat.check(lib.cc_hip_components_last_pass_ran(None, ctypes.byref(ran)), "last pass")
        D.close()
        check_labels(G, got)
        seen.add(ran.value)
    assert seen == {0, 1}


def test_solve_on_caller_streams(gpu):
    """Solves on streams the caller creates and destroys in turn: each
    stream's scratch state is made on first use and freed with the stream,
    and the labels are exact every time."""
    lib = nat.gpu_lib()
    G = cm.rmat(14, 16, seed=9)
    D = cm.DeviceGraph(G)
    for _ in range(3):
        s = ctypes.c_void_p()
        nat.check(lib.cc_hip_stream_create(ctypes.byref(s)), "stream create")
        try:
            got = D.solve(s)
        finally:
            nat.check(lib.cc_hip_stream_destroy(s), "stream destroy")
        check_labels(G, got)
    assert np.array_equal(D.solve(), got)
    D.close()


def test_low_first_plan_on_concurrent_streams(gpu):
    """Two dense graphs of 2^23 vertices (the low-first hook from the CSR,
    queued walks) solved at the same time on two caller streams, three times
    each with no wait in between: each stream keeps its own giant root, map
    and workspace, so both label arrays equal the oracle's."""
    from cc_mpi_amd import devgen
    lib = nat.gpu_lib()
    graphs = [devgen.rmat(23, 16, seed=61), devgen.chung_lu(1 << 23, 245000000, 2.1, seed=62)]
    for D in graphs:
        assert lib.cc_hip_components_plan_get(D.n, D.m, b"csr_hook") == 1
        assert lib.cc_hip_components_plan_get(D.n, D.m, b"low_shift") == 4
    streams = [ctypes.c_void_p(), ctypes.c_void_p()]
    for s in streams:
        nat.check(lib.cc_hip_stream_create(ctypes.byref(s)), "stream create")
    try:
        for _ in range(3):
            for D, s in zip(graphs, streams):
                D.solve_async(s)
        got = [D.labels.download(np.empty(D.n, dtype=np.uint32), s)
               for D, s in zip(graphs, streams)]
    finally:
        for s in streams:
            nat.check(lib.cc_hip_stream_destroy(s), "stream destroy")
    for D, g in zip(graphs, got):
        check_labels(D.to_host(), g)
        D.close()


def test_device_resident_repeat(gpu):
    # same labels on repeated solves of a resident graph, and equal to the oracle
    G = cm.rmat(16, 16, seed=5)
    D = cm.DeviceGraph(G)
    first = D.solve()
    check_labels(G, first)
    for _ in range(5):
        assert np.array_equal(D.solve(), first)
    D.close()


@pytest.mark.parametrize("make,tiles", [
    (lambda: lattice3d(256, 0.5, seed=61), 1),          # a third of the entries cross tiles
    (lambda: banded(1 << 21, 4, 1 << 13, seed=62), 1),   # local, low degree
    (lambda: banded(1 << 21, 8, 1 << 20, seed=63), 0),   # edges spanning 2^20 ids
    (lambda: cm.grid(1448, 1448, 0.7, seed=64), 0),      # permuted ids
], ids=["lattice3d", "banded_local", "banded_wide", "grid_permuted"])
def test_probe_choice_on_local_and_nonlocal_graphs(gpu, make, tiles):
    """Sparse graphs of 2^20+ vertices: the probe on the device picks the
    vertex-tile plan where ids are local and the sampled plan elsewhere;
    labels are exact either way."""
    lib = nat.gpu_lib()
    G = make()
    assert G.n >= 1 << 20
    assert lib.cc_hip_components_plan_get(G.n, G.m, b"tiles") == 2
    D = cm.DeviceGraph(G)
    got = D.solve()
    took = ctypes.c_int(-1)
    nat.check(lib.cc_hip_components_tiles_ran(None, ctypes.byref(took)), "tiles ran")
    D.close()
    assert took.value == tiles
    check_labels(G, got)


def test_rmat22_full_size(gpu):
    # R-MAT scale 22, edge factor 16, (0.57,0.19,0.19,0.05), permuted ids
    G = cm.rmat(22, 16, seed=1002)
    D = cm.DeviceGraph(G)
    got = D.solve()
    D.close()
    want = check_labels(G, got)
    # component count and size histogram agree
    roots = np.flatnonzero(want == np.arange(G.n))
    assert np.count_nonzero(got == np.arange(G.n)) == roots.size
    assert np.array_equal(np.bincount(got, minlength=G.n)[roots],
                          np.bincount(want, minlength=G.n)[roots])


PATHS = [  # every code path of the solve: (knobs) -> same labels
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=1),
    dict(sample_cache=0, fused_pair=0, hook_ilp=2, hook_mode=1),
    dict(sample_cache=0, fused_pair=0, hook_ilp=4, hook_mode=1),
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=0),
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=2),
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=3),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=1),
    dict(sample_cache=1, fused_pair=0, hook_ilp=2, hook_mode=1),
    dict(sample_cache=1, fused_pair=0, hook_ilp=4, hook_mode=1),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=0),
    dict(sample_cache=1, fused_pair=0, hook_ilp=4, hook_mode=1, giant_early=0),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=0, giant_early=0),
    dict(sample_cache=1, fused_pair=0, hook_ilp=2, hook_mode=1, giant_early=1, sampled_neighbours=4),
    dict(sample_cache=1, fused_pair=1, hook_ilp=1, hook_mode=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=2, hook_mode=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=1, hook_mode=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=2, hook_mode=1, sampled_neighbours=1),
    dict(sample_cache=1, fused_pair=0, hook_ilp=4, hook_mode=1, sampled_neighbours=3),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=1, sampled_neighbours=4),
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=4),
    dict(sample_cache=0, fused_pair=0, hook_ilp=2, hook_mode=4),
    dict(sample_cache=0, fused_pair=0, hook_ilp=4, hook_mode=4),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=4),
    dict(sample_cache=1, fused_pair=0, hook_ilp=2, hook_mode=4, giant_early=0),
    dict(sample_cache=1, fused_pair=0, hook_ilp=4, hook_mode=4),
    dict(sample_cache=1, fused_pair=1, hook_ilp=1, hook_mode=4, halve=0, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=2, hook_mode=4, halve=0, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=4, hook_mode=4, halve=0, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=1, hook_mode=4, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=2, hook_mode=4, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=4, hook_mode=4, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=1, hook_mode=1, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=2, hook_mode=1, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=4, hook_mode=1, halve=1, rem=0),
    dict(sample_cache=1, fused_pair=1, hook_ilp=4, hook_mode=4, halve=1, rem=0, blocks_per_cu=1),
    dict(sample_cache=1, fused_pair=1, hook_mode=4, rem=1, rem_splice=0),
    dict(sample_cache=1, fused_pair=1, hook_mode=4, rem=1, rem_splice=1),
    dict(sample_cache=1, fused_pair=1, hook_mode=1, rem=1, rem_splice=1),
    dict(sample_cache=1, fused_pair=1, hook_mode=4, rem=1, rem_splice=1, blocks_per_cu=1),
    dict(sample_cache=1, fused_pair=1, hook_mode=4, rem=1, rem_splice=2),
    dict(sample_cache=1, fused_pair=1, hook_mode=1, rem=1, rem_splice=2, blocks_per_cu=2),
    dict(sample_cache=1, fused_pair=1, hook_mode=4, rem=0, halve=0),
    dict(stream_ilp=2), dict(stream_ilp=4, sample_cache=0),
    # non-temporal streaming and the init pass's rows per thread
    dict(nt=1), dict(nt=2), dict(nt=3), dict(nt=7, init_ilp=2), dict(nt=5, init_ilp=4),
    dict(nt=4, rem=1, rem_splice=1), dict(nt=4, rem=1, rem_splice=2),
    dict(nt=5, rem=0, halve=1, hook_ilp=1), dict(nt=5, rem=0, halve=1, hook_ilp=2),
    dict(nt=5, rem=0, halve=1, hook_ilp=4, init_ilp=4), dict(nt=0, init_ilp=4),
    # frontier sweep: chunks per wave step, with every hook mode
    dict(frontier_chunks=1), dict(frontier_chunks=2), dict(frontier_chunks=4),
    dict(frontier_chunks=4, hook_mode=0), dict(frontier_chunks=2, hook_mode=1),
    dict(frontier_chunks=4, hook_mode=2, sample_cache=0, fused_pair=0),
    dict(frontier_chunks=4, hook_mode=3, sample_cache=0, fused_pair=0),
    # the one-pass hooks that read the rows themselves after an identity init
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1), dict(csr_hook=1, rem=0, halve=1, hook_ilp=2),
    dict(csr_hook=1, rem=0, halve=0, hook_ilp=4), dict(csr_hook=1, rem=0, halve=1, hook_ilp=4),
    dict(csr_hook=1, rem=1, rem_splice=0), dict(csr_hook=1, rem=1, rem_splice=1),
    dict(csr_hook=1, rem=1, rem_splice=2), dict(csr_hook=0, rem=0, halve=1, hook_ilp=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, blocks_per_cu=1),
    # linked vertices skip neighbour 1 (1), or also take neighbour 0's parent
    # (2); local minima unite with minima_entries entries; off (0) where the
    # plan would skip
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2),
    dict(csr_hook=1, rem=0, halve=0, hook_ilp=2, skip_linked=2),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=4, skip_linked=2, blocks_per_cu=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, minima_entries=3),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=2, skip_linked=1, minima_entries=8),
    # the low-first hook (the one-hop skipping hook from the CSR, one root
    # per thread): off, half of the ids first, the smallest bound (64 ids,
    # grid-stride loops of many steps), with more entries for local minima
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=0),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=8, blocks_per_cu=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=2, minima_entries=3,
         stream_blocks_per_cu=1),
    # ... split into own-entry links and a dense pass over the walk list
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=4, split_walks=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=1, split_walks=1,
         blocks_per_cu=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=8, split_walks=0),
    dict(rem=1, rem_splice=0, skip_linked=1), dict(rem=1, rem_splice=1, skip_linked=2),
    dict(rem=1, rem_splice=2, skip_linked=2, nt=4), dict(rem=1, rem_splice=1, skip_linked=1, nt=4),
    dict(rem=1, rem_splice=0, skip_linked=2, minima_entries=4),
    dict(rem=1, rem_splice=1, skip_linked=1, minima_entries=8, blocks_per_cu=1),
    dict(rem=0, halve=1, hook_ilp=1, skip_linked=2), dict(rem=0, halve=1, hook_ilp=2, skip_linked=1),
    dict(rem=0, halve=0, hook_ilp=4, skip_linked=2), dict(rem=0, halve=1, hook_ilp=2, skip_linked=2, nt=5),
    dict(skip_linked=2, stream_ilp=2), dict(skip_linked=0),
    # vertex tiles: whole plan forced on every graph, at several tile sizes,
    # with pair segments small enough that tiles overflow and read their rows
    # again; and forced off
    dict(tiles=1), dict(tiles=1, tile_shift=6), dict(tiles=1, tile_shift=9, tile_pairs=1),
    dict(tiles=1, tile_shift=12, tile_pairs=64), dict(tiles=1, tile_shift=7, stream_blocks_per_cu=1),
    dict(tiles=1, tile_shift=16), dict(tiles=1, tile_shift=16, tile_pairs=3),
    dict(tiles=0),
]
DEFAULTS = dict(sample_cache=1, fused_pair=-1, hook_ilp=0, hook_mode=4, sampled_neighbours=2,
                stream_ilp=1, blocks_per_cu=0, giant_early=1, halve=-1, rem=-1, rem_splice=-1,
                nt=-1, init_ilp=0, frontier_chunks=0, csr_hook=-1, skip_linked=-1, minima_entries=0,
                tiles=-1, tile_shift=0, tile_pairs=0, stream_blocks_per_cu=8, low_shift=-1,
                split_walks=-1)


@pytest.mark.parametrize("knobs", PATHS, ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()))
def test_every_solve_path(gpu, knobs):
    """Labels are a function of the graph alone, whichever kernels run."""
    lib = nat.gpu_lib()
    graphs = [cm.rmat(15, 16, seed=31), cm.erdos_renyi(50000, 3.0, seed=32),
              cm.grid(120, 150, 0.6, seed=33), disjoint_cliques(200, 5, seed=34),
              cm.chung_lu(1 << 16, 1 << 19, 2.1, seed=35),
              cm.grid(130, 170, 0.6, seed=36, permute=False)]
    try:
        for k, v in knobs.items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)
        for G in graphs:
            D = cm.DeviceGraph(G)
            got = D.solve()
            again = D.solve()
            D.close()
            check_labels(G, got)
            assert np.array_equal(got, again)
    finally:
        for k, v in DEFAULTS.items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)



@pytest.mark.parametrize("knobs", [
    dict(), dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=1, skip_linked=2, low_shift=2, split_walks=1),
    dict(csr_hook=1, rem=0, halve=1, hook_ilp=2, skip_linked=0),
    dict(rem=1, rem_splice=1, skip_linked=2), dict(rem=1, rem_splice=2, skip_linked=0),
    dict(rem=0, halve=1, hook_ilp=2, skip_linked=2),
    dict(sample_cache=1, fused_pair=0, hook_ilp=2, hook_mode=1, giant_early=1),
    dict(sample_cache=1, fused_pair=0, hook_ilp=1, hook_mode=4, giant_early=1, sampled_neighbours=3),
    dict(sample_cache=0, fused_pair=0, hook_ilp=1, hook_mode=3),
    dict(tiles=1, tile_shift=7), dict(tiles=1, tile_shift=6, tile_pairs=1)],
    ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()) or "default")
def test_one_directional_entries_with_symmetric_off(gpu, knobs):
    """With the knob symmetric = 0 the solve unites every stored entry, so a
    CSR holding some edges in one direction only gets its weakly connected
    components (the reference's propagation has no schedule-independent
    answer there; tests/test_oracle.py::test_reference_needs_both_directions)."""
    lib = nat.gpu_lib()
    graphs = [one_directional(cm.rmat(13, 16, seed=51), 0.2, 1),
              one_directional(cm.grid(90, 110, 0.7, seed=52), 0.1, 2),
              one_directional(cm.grid(80, 120, 0.7, seed=55, permute=False), 0.2, 5),
              one_directional(cm.erdos_renyi(30000, 3.0, seed=53), 0.3, 3),
              one_directional(cm.chung_lu(1 << 15, 1 << 18, 2.1, seed=54), 0.0, 4)]
    try:
        for k, v in dict(knobs, symmetric=0).items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)
        for G in graphs:
            D = cm.DeviceGraph(G)
            got = D.solve()
            D.close()
            want = oracle.cc_min_label(G.row_ptr, G.col_idx)
            bad = np.flatnonzero(got != want)
            assert bad.size == 0, f"{bad.size} labels differ, first {bad[:5]}"
    finally:
        for k, v in dict(DEFAULTS, symmetric=1).items():
            nat.check(lib.cc_hip_components_set(k.encode(), v), k)


def test_drop_ins_from_four_host_threads(gpu):
    """compute_connected_components_pthreads, boundary_edges_sort and
    boundary_edges_dedup called from 4 host threads at once (ctypes releases
    the interpreter lock during the calls), as the reference's stateless
    functions may be (src/cc_pthreads.c:152-229): every result equals the
    oracle's."""
    import threading
    from cc_mpi_amd import boundary as bd
    rng = np.random.default_rng(77)
    graphs = [cm.rmat(14, 16, seed=60), cm.grid(300, 300, 0.7, seed=61),
              cm.erdos_renyi(200_000, 3.0, seed=62), cm.rmat(15, 8, seed=63)]
    want = [oracle.cc_min_label(G.row_ptr, G.col_idx) for G in graphs]
    keys = [rng.integers(0, 1 << 50, int(n), dtype=np.uint64) // np.uint64(7 ** i)
            for i, n in enumerate((1_500_000, 700_000, 2_100_000, 300_000))]
    want_sorted = [oracle.sort_u64(k) for k in keys]
    want_dedup = [oracle.dedup_u64(k) for k in want_sorted]
    errors = []
    start = threading.Barrier(4)

    def worker(i):
        try:
            start.wait()
            for rep in range(5):
                got = cm.compute_connected_components_pthreads(graphs[i])
                if not np.array_equal(got, want[i]):
                    errors.append(f"thread {i} rep {rep}: labels differ")
                k = keys[i].copy()
                bd.boundary_edges_sort(k)
                if not np.array_equal(k, want_sorted[i]):
                    errors.append(f"thread {i} rep {rep}: sort differs")
                cnt = bd.boundary_edges_dedup(k)
                if not np.array_equal(k[:cnt], want_dedup[i]):
                    errors.append(f"thread {i} rep {rep}: dedup differs")
        except Exception as e:  # noqa: BLE001
            errors.append(f"thread {i}: {e!r}")

    threads = [threading.Thread(target=worker, args=(i,)) for i in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
