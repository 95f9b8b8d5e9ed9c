"""Host logic of bench.py that needs no GPU: the frontier statistics that
the algorithmic-byte model of the solve's kernels is built from, checked
against a union-find restatement written here, and the one-GPU rule of
the command line."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import cc_mpi_amd as cm  # noqa: E402


def _brute_stats(G, skip=False):
    """deg_gt0/deg_gt1/up0 by a loop over the vertices; the sampled forest
    (every vertex joined with its neighbours 0 and 1) by a union-find; the
    giant tree is the largest one (the graphs below make it hold most
    vertices, so the sampled mode is it); frontier = vertices outside it
    with more than two neighbours, and their entries past the first two.
    skip: only local minima (neighbour 0 above them) join neighbour 1, and
    the frontier counts entries past the first."""
    n = G.n
    rp = G.row_ptr.astype(np.int64)
    ci = G.col_idx.astype(np.int64)
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    d0 = d1 = up0 = m1 = 0
    for v in range(n):
        deg = rp[v + 1] - rp[v]
        d0 += deg > 0
        d1 += deg > 1
        minimum = deg > 0 and ci[rp[v]] > v
        up0 += minimum
        m1 += minimum and deg > 1
        for r in range(min(2 if (minimum or not skip) else 1, deg)):
            a, b = find(v), find(int(ci[rp[v] + r]))
            if a != b:
                parent[max(a, b)] = min(a, b)
    roots = [find(v) for v in range(n)]
    vals, counts = np.unique(roots, return_counts=True)
    giant = vals[np.argmax(counts)]
    fv = fe = 0
    past = 1 if skip else 2
    for v in range(n):
        deg = rp[v + 1] - rp[v]
        if roots[v] != giant and deg > past:
            fv += 1
            fe += deg - past
    return dict(deg_gt0=d0, deg_gt1=d1, up0=up0, minima_gt1=m1, frontier_vertices=fv,
                frontier_edges=fe)


def _two_part_graph(seed):
    """An R-MAT graph (its sampled forest has one dominant tree) plus a
    disjoint clique block and isolated vertices: vertices that stay outside
    the giant tree, some with more than two neighbours."""
    G = cm.rmat(12, 8, seed=seed)
    n0 = G.n
    k = 9  # a 9-clique: degree 8, so 6 entries past the first two each
    rows = [list(G.col_idx[G.row_ptr[v]:G.row_ptr[v + 1]]) for v in range(n0)]
    for i in range(k):
        rows.append([n0 + j for j in range(k) if j != i])
    rows += [[], []]
    row_ptr = np.zeros(len(rows) + 1, dtype=np.uint64)
    row_ptr[1:] = np.cumsum([len(r) for r in rows])
    col_idx = np.array([c for r in rows for c in r], dtype=np.uint32)
    return _graph(len(rows), row_ptr, col_idx)


def _graph(n, row_ptr, col_idx):
    class G:  # the fields bench.frontier_stats reads
        pass
    g = G()
    g.n, g.row_ptr, g.col_idx = n, row_ptr, col_idx
    return g


def test_frontier_stats_match_a_union_find_restatement():
    for seed in (3, 4):
        for skip in (False, True):
            G = _two_part_graph(seed)
            got = bench.frontier_stats(G, two_pass=False, skip=skip)
            want = _brute_stats(G, skip)
            for k, v in want.items():
                assert got[k] == v, (seed, skip, k, got[k], v)
            assert want["frontier_vertices"] >= 9  # the clique block is outside the giant
            assert 0.5 < got["giant_fraction"] < 1.0


def test_low_split_counts_match_a_loop():
    """The low-first hook's byte-model inputs: the counts split at the bound
    L, and the linked vertices at or above L whose neighbour 0 lies below L
    in the giant tree of the sampled forest (the skipping hook's edge set),
    by a loop over the vertices and a union-find."""
    for seed, shift in ((5, 2), (6, 4)):
        G = _two_part_graph(seed)
        n = G.n
        L = bench.low_bound(n, shift)
        assert L % 64 == 0 and 0 < L < n
        got = bench.frontier_stats(G, two_pass=False, skip=True, low=L)
        rp = G.row_ptr.astype(np.int64)
        ci = G.col_idx.astype(np.int64)
        parent = list(range(n))

        def find(x):
            while parent[x] != x:
                x = parent[x]
            return x

        want = {f"{k}_{t}": 0 for k in ("deg_gt0", "up0", "minima_gt1") for t in ("lo", "hi")}
        for v in range(n):
            deg = rp[v + 1] - rp[v]
            tag = "lo" if v < L else "hi"
            minimum = deg > 0 and ci[rp[v]] > v
            want[f"deg_gt0_{tag}"] += deg > 0
            want[f"up0_{tag}"] += minimum
            want[f"minima_gt1_{tag}"] += minimum and deg > 1
            for r in range(min(2 if minimum else 1, deg)):
                a, b = find(v), find(int(ci[rp[v] + r]))
                if a != b:
                    parent[max(a, b)] = min(a, b)
        roots = np.array([find(v) for v in range(n)])
        vals, counts = np.unique(roots, return_counts=True)
        giant = vals[np.argmax(counts)]
        marked = 0
        for v in range(L, n):
            if rp[v + 1] > rp[v]:
                w0 = int(ci[rp[v]])
                marked += w0 < v and w0 < L and roots[w0] == giant
        want["marked_hi"] = marked
        for k, v in want.items():
            assert got[k] == v, (seed, k, got[k], v)
        assert got["low_bound"] == L and marked > 0


def test_low_bound_rule():
    # ids below n >> shift rounded up to whole 64-id map words, at least 64,
    # none when that reaches n
    assert bench.low_bound(65601024, 4) == 4100096
    assert bench.low_bound(210, 8) == 64
    assert bench.low_bound(210, 1) == 128
    assert bench.low_bound(64, 1) == 0
    assert bench.low_bound(50, 8) == 0
    assert bench.low_bound(1 << 20, 0) == 0


def test_tile_stats_match_a_union_find_restatement():
    """The vertex-tile byte model's inputs (pairs to smaller tiles, local
    roots, linked local roots, vertices whose local root is not final) on
    small lattices numbered row by row and permuted, against a plain
    union-find over the entries inside each tile."""
    import oracle
    for permute, shift in ((False, 7), (True, 6), (False, 9)):
        G = cm.grid(40, 64, 0.7, seed=3, permute=permute)
        lab = oracle.cc_min_label(G.row_ptr, G.col_idx)
        st = bench.tile_stats(G, shift, lab)
        n, T = G.n, 1 << shift
        par = list(range(n))

        def find(x):
            while par[x] != x:
                x = par[x]
            return x

        cross = 0
        for v in range(n):
            for w in G.col_idx[G.row_ptr[v]:G.row_ptr[v + 1]].tolist():
                cross += w // T < v // T
                if w // T == v // T and w != v:
                    a, b = find(v), find(w)
                    if a != b:
                        par[max(a, b)] = min(a, b)
        local = [find(v) for v in range(n)]
        roots = sum(local[v] == v for v in range(n))
        assert st["cross_pairs"] == cross
        assert st["local_roots"] == roots
        assert st["linked_local_roots"] == roots - int(np.count_nonzero(lab == np.arange(n)))
        assert st["moved"] == sum(local[v] != v and lab[v] != local[v] for v in range(n))


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stdout + r.stderr
