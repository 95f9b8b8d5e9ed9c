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

The low-first hook (k_hook_pair_csr with Map) is modelled too: the ids
below a bound L run their hooks to the end first; then g0 is the most
frequent root among them and the map marks the ids below L whose root is
g0; then the other ids hook, a vertex whose neighbour 0 is marked taking
g0 in place of the value read at step 1, and a failed link with a marked
neighbour 0 uniting from g0.  With failed_w1 off, a vertex whose CAS
failed unites with neighbour 0 only: correct too, but the device keeps
both unions for such vertices (uniting only one left more trees apart and
made the frontier sweep slower).
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


def hook(mem, v, row, rng, seen, failed_w1=True, marked=None, g0=None):
    """Generator: one yield per step of vertex v's hook.  marked(w): the
    low-id map's bit for w (None: no map)."""
    if not row:
        return
    w0 = row[0]
    w1 = row[1] if len(row) > 1 else None
    linked = False
    on_map = w0 < v and marked is not None and marked(w0)
    if w0 < v:
        if on_map:
            q = g0
        else:
            h = mem.hist[w0]
            i = rng.randint(seen.get(w0, 0), len(h) - 1)
            seen[w0] = i
            q = min(h[i], w0)
        yield
        if mem.latest(v) == v:  # CAS(p[v], v, q)
            assert q < v
            mem.write(v, q)
            linked = True
        yield
    if not linked and w0 != v:
        unite(mem, v, g0 if on_map else w0)
        yield
        if w1 is not None and w1 != v and (failed_w1 or w0 > v):
            unite(mem, v, w1)
            yield


def interleave(gens, ids, rng):
    live = list(ids)
    while live:
        i = rng.choice(live)
        try:
            next(gens[i])
        except StopIteration:
            live.remove(i)


def run(n, rows, seed, first=1, failed_w1=True, low=0, plant_all_marked=False):
    rng = random.Random(seed)
    mem = Mem(n)
    seen = [dict() for _ in range(n)]
    if low:
        # the low pass, then the map over [0, low), then the rest
        L = min(low, n)
        gens = {v: hook(mem, v, rows[v], rng, seen[v], failed_w1) for v in range(L)}
        interleave(gens, range(L), rng)
        lows = [root(mem, v) for v in range(L)]
        counts = {}
        for r in lows:
            counts[r] = counts.get(r, 0) + 1
        g0 = min(counts, key=lambda r: (-counts[r], r)) if lows else None
        bits = [plant_all_marked or r == g0 for r in lows]

        def marked(w):
            return w < L and bits[w]
        gens = {v: hook(mem, v, rows[v], rng, seen[v], failed_w1, marked, g0)
                for v in range(L, n)}
        interleave(gens, range(L, n), rng)
    else:
        gens = {v: hook(mem, v, rows[v], rng, seen[v], failed_w1) for v in range(n)}
        interleave(gens, range(n), rng)
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


def test_failed_links_unite_neighbour_0_only():
    """A vertex whose own-entry CAS failed (another union linked it first)
    unites with neighbour 0 alone, like a vertex whose link succeeded: the
    sweep from entry 1 still covers the rest."""
    for trial in range(600):
        rng = random.Random(20000 + trial)
        n, rows = random_rows(rng)
        assert run(n, rows, trial, failed_w1=False), f"trial {trial}"


@pytest.mark.parametrize("failed_w1", [True, False])
def test_low_first_model_gives_component_minima(failed_w1):
    for trial in range(600):
        rng = random.Random(30000 + trial)
        n, rows = random_rows(rng)
        low = (n // 2, n // 4, 8, 1)[trial % 4]
        assert run(n, rows, trial, failed_w1=failed_w1, low=low), f"trial {trial}"


def test_low_first_model_catches_a_map_marking_every_id():
    """Planted mistake: a map that marks every low id (not only those under
    g0) links vertices of other components under g0."""
    failures = 0
    for trial in range(600):
        rng = random.Random(40000 + trial)
        n, rows = random_rows(rng)
        failures += not run(n, rows, trial, failed_w1=False, low=max(1, n // 2),
                            plant_all_marked=True)
    assert failures > 0


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
        assert run(n, rows, trial, failed_w1=False, low=n // 4), f"trial {trial}"
