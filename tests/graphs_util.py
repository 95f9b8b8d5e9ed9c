"""Small hand-built graphs for edge cases (shared by CPU and GPU tests)."""
import numpy as np

from cc_mpi_amd.graphs import Graph, from_pairs


def csr_raw(n, adj):
    """CSR exactly as given (no symmetrisation/dedup): adj is a list of rows."""
    rp = np.zeros(n + 1, dtype=np.uint64)
    for i, row in enumerate(adj):
        rp[i + 1] = rp[i] + len(row)
    ci = np.array([c for row in adj for c in row], dtype=np.uint32)
    return Graph(rp, ci, name="raw")


def path(n, perm_seed=None):
    src = np.arange(n - 1, dtype=np.uint32)
    dst = src + 1
    perm = None
    if perm_seed is not None:
        perm = np.random.default_rng(perm_seed).permutation(n).astype(np.uint32)
    return from_pairs(n, src, dst, perm=perm, name=f"path-{n}")


def star(n, centre=0):
    leaves = np.array([v for v in range(n) if v != centre], dtype=np.uint32)
    return from_pairs(n, np.full(leaves.size, centre, dtype=np.uint32), leaves, name=f"star-{n}")


def complete(n):
    s, d = np.triu_indices(n, 1)
    return from_pairs(n, s.astype(np.uint32), d.astype(np.uint32), name=f"K{n}")


def disjoint_cliques(k, size, seed=0):
    """k cliques of `size` vertices with ids shuffled over [0, k*size)."""
    n = k * size
    perm = np.random.default_rng(seed).permutation(n).astype(np.uint32)
    s, d = np.triu_indices(size, 1)
    src = np.concatenate([s + i * size for i in range(k)]).astype(np.uint32)
    dst = np.concatenate([d + i * size for i in range(k)]).astype(np.uint32)
    return from_pairs(n, src, dst, perm=perm, name=f"{k}xK{size}")


def lattice3d(side, keep, seed):
    """side^3 lattice numbered row by row (id = x + side*y + side^2*z), each
    of its 6-neighbour edges kept with probability `keep`."""
    rng = np.random.default_rng(seed)
    n = side ** 3
    v = np.arange(n, dtype=np.int64)
    coord = (v % side, (v // side) % side, v // (side * side))
    src, dst = [], []
    for c, step in zip(coord, (1, side, side * side)):
        sel = v[(c + 1 < side) & (rng.random(n) < keep)]
        src.append(sel)
        dst.append(sel + step)
    return from_pairs(n, np.concatenate(src), np.concatenate(dst), name=f"lattice3d-{side}")


def banded(n, degree, width, seed):
    """n*degree/2 random edges, each from a uniform vertex to one up to
    `width` ids above it (dropped past n)."""
    rng = np.random.default_rng(seed)
    m = n * degree // 2
    a = rng.integers(0, n, m)
    b = a + rng.integers(1, width, m)
    ok = b < n
    return from_pairs(n, a[ok], b[ok], name=f"banded-{n}-d{degree}-w{width}")


def bfs_min_labels(G):
    """Brute force: BFS from each unvisited vertex in id order (so the
    starting vertex is the component minimum)."""
    n = G.n
    lab = np.full(n, 0xFFFFFFFF, dtype=np.uint32)
    rp, ci = G.row_ptr, G.col_idx
    for s in range(n):
        if lab[s] != 0xFFFFFFFF:
            continue
        lab[s] = s
        stack = [s]
        while stack:
            u = stack.pop()
            for j in range(int(rp[u]), int(rp[u + 1])):
                w = int(ci[j])
                if w < n and lab[w] == 0xFFFFFFFF:
                    lab[w] = s
                    stack.append(w)
    return lab


def one_directional(G, keep_both, seed):
    """G with only one direction kept of each edge, except a fraction
    `keep_both` stored both ways: a CSR whose weakly connected components
    are G's components, but whose rows no longer see every edge."""
    rng = np.random.default_rng(seed)
    deg = np.diff(G.row_ptr).astype(np.int64)
    src = np.repeat(np.arange(G.n, dtype=np.int64), deg)
    dst = G.col_idx.astype(np.int64)
    lo, hi = np.minimum(src, dst), np.maximum(src, dst)
    # one draw per undirected edge, shared by its two entries
    h = (lo * 0x9E3779B1 + hi * 0x85EBCA77 + seed) & 0xFFFFFFFF
    both = (h % 1000) < keep_both * 1000
    forward = ((h >> 10) & 1) == 1
    keep = both | (forward == (src < dst))
    rows = [[] for _ in range(G.n)]
    for s, d in zip(src[keep].tolist(), dst[keep].tolist()):
        rows[s].append(d)
    for r in rows:
        rng.shuffle(r)
    return csr_raw(G.n, rows)
