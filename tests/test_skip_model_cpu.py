"""Model of the sampled hook with skipping (k_hook_pair_csr and k_hook_pair
with Skip1 = 2, k_hook_rem with Skip1 in components.hip) followed by the
frontier sweep, run on the CPU.

Each vertex's hook is a small program whose steps interleave with the
other vertices' in random order:
  1. read p[w0] for the one-hop link (a plain read: any value from the
     address's history no older than the last one this vertex saw there);
  2. CAS p[v] from v to min(that value, w0), if w0 is below v;
  3. unions, each one step on the latest values (the concurrent union
     itself is modelled step by step in test_rem_model_cpu.py): with w0 and
     w1 for a local minimum (w0 above v) or when the CAS failed, none else.
Then the giant root is the most frequent root, and every vertex outside
its tree unites with its entries from 1 on.  The roots must be each
component's minimum.  Rows are given in random order (the device needs
no row order), and some graphs have isolated vertices and duplicates.
A planted mistake, starting the sweep at entry 2 as the plans without
skipping do, must fail on some graphs.
"""
import random

import pytest


class Mem:
    def __init__(self, n):
        self.hist = [[v] for v in range(n)]

    def latest(self, a):
        return self.hist[a][-1]

    def write(self, a, v):
        if v != self.hist[a][-1]:
            self.hist[a].append(v)


def root(mem, x):
    while mem.latest(x) != x:
        assert mem.latest(x) < x  # p[v] <= v throughout
        x = mem.latest(x)
    return x


def unite(mem, a, b):
    ra, rb = root(mem, a), root(mem, b)
    if ra != rb:
        mem.write(max(ra, rb), min(ra, rb))


def hook(mem, v, row, rng, seen):
    """Generator: one yield per step of vertex v's hook."""
    if not row:
        return
    w0 = row[0]
    w1 = row[1] if len(row) > 1 else None
    linked = False
    if w0 < v:
        h = mem.hist[w0]
        i = rng.randint(seen.get(w0, 0), len(h) - 1)
        seen[w0] = i
        q = h[i]
        yield
        if mem.latest(v) == v:  # CAS(p[v], v, min(q, w0))
            mem.write(v, min(q, w0))
            linked = True
        yield
    if not linked and w0 != v:
        unite(mem, v, w0)
        yield
        if w1 is not None and w1 != v:
            unite(mem, v, w1)
            yield


def run(n, rows, seed, first=1):
    rng = random.Random(seed)
    mem = Mem(n)
    seen = [dict() for _ in range(n)]
    gens = [hook(mem, v, rows[v], rng, seen[v]) for v in range(n)]
    live = list(range(n))
    while live:
        i = rng.choice(live)
        try:
            next(gens[i])
        except StopIteration:
            live.remove(i)
    roots = [root(mem, v) for v in range(n)]
    counts = {}
    for r in roots:
        counts[r] = counts.get(r, 0) + 1
    giant = min(counts, key=lambda r: (-counts[r], r))
    outside = [v for v in range(n) if roots[v] != giant]
    rng.shuffle(outside)
    for v in outside:
        for w in rows[v][first:]:
            unite(mem, v, w)
    got = [root(mem, v) for v in range(n)]
    par = list(range(n))

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    for v in range(n):
        for w in rows[v]:
            ra, rb = find(v), find(w)
            if ra != rb:
                par[max(ra, rb)] = min(ra, rb)
    return got == [find(v) for v in range(n)]


def random_rows(rng):
    """A symmetric multigraph with self-loop-free rows in random order;
    some graphs are a few dense clusters joined by single edges (so that
    local minima's first entries often stay inside their cluster)."""
    n = rng.randint(2, 80)
    edges = []
    if rng.random() < 0.5:
        edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 3 * n))]
    else:
        k = rng.randint(1, 6)
        for v in range(n):
            c = v % k
            for _ in range(rng.randint(0, 4)):
                w = rng.randrange(n)
                if w % k == c:
                    edges.append((v, w))
        for _ in range(rng.randint(0, k)):
            edges.append((rng.randrange(n), rng.randrange(n)))
    rows = [[] for _ in range(n)]
    for a, b in edges:
        if a != b:
            rows[a].append(b)
            rows[b].append(a)
    for r in rows:
        if rng.random() < 0.7:
            r.sort()
        else:
            rng.shuffle(r)
    return n, rows


def test_skip_model_gives_component_minima():
    for trial in range(600):
        rng = random.Random(trial)
        n, rows = random_rows(rng)
        assert run(n, rows, trial), f"trial {trial}"


def test_skip_model_catches_sweep_from_entry_2():
    failures = 0
    for trial in range(600):
        rng = random.Random(9000 + trial)
        n, rows = random_rows(rng)
        failures += not run(n, rows, trial, first=2)
    assert failures > 0


@pytest.mark.parametrize("n,k", [(64, 1), (90, 3), (40, 8)])
def test_skip_model_on_cliques_and_paths(n, k):
    """Disjoint cliques and paths with permuted ids: the minima's links stay
    inside their own component, and the sweep must finish every one."""
    for trial in range(40):
        rng = random.Random(100 * n + trial)
        perm = list(range(n))
        rng.shuffle(perm)
        rows = [[] for _ in range(n)]
        for v in range(n):
            c = v % k
            members = [u for u in range(c, n, k)]
            if trial % 2:  # cliques
                rows[perm[v]] = [perm[u] for u in members if u != v]
            else:  # paths
                i = members.index(v)
                rows[perm[v]] = [perm[u] for u in members[max(0, i - 1):i + 2] if u != v]
        for r in rows:
            r.sort()
        assert run(n, rows, trial), f"trial {trial}"
