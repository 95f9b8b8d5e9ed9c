"""Step-level model of the concurrent Rem's union used by the sampled hook
(k_hook_rem in components.hip, rem_unite in batch.hip), run on the CPU.

Simulated threads interleave one memory operation at a time in random
order.  A plain read may return any value from the address's history that
is no older than the last value that thread saw there (a stale but
monotone cache, as an L1 line can be); the L2-confirm read and the atomics
(CAS link, atomic-min splice) act on the latest value.  Every run must end
within a step bound, and the final forest's roots must be each component's
minimum id.  A planted mistake (linking without the CAS check) must fail,
which shows the model can catch a lost link.
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


def rem_unite(mem, seen, rx, ry, splice, rng, planted=False):
    """Generator: one yield per memory operation of the device loop.
    splice: 0 never, 1 every step, 2 only the first vertex's own entry."""
    v = rx
    def plain(a):
        h = mem.hist[a]
        i = rng.randint(seen.get(a, 0), len(h) - 1)
        seen[a] = i
        return h[i]

    px = plain(rx)
    yield
    py = plain(ry)
    yield
    while px != py:
        if px < py:
            rx, ry, px, py = ry, rx, py, px
        assert px > py
        if rx == px:  # looks like a root
            cur = mem.latest(rx)  # L2 confirm
            yield
            if cur != rx:
                px = cur
                continue
            old = mem.latest(rx)  # CAS(p[rx], rx, py)
            if old == rx or planted:
                mem.write(rx, py)
                old = rx
            yield
            if old == rx:
                return
            px = old
            continue
        if splice == 1 or (splice == 2 and rx == v):  # atomic min of p[rx] to py
            if py < mem.latest(rx):
                mem.write(rx, py)
            yield
        rx = px
        px = plain(rx)
        yield


def run(n, edges, threads, splice, seed, planted=False):
    rng = random.Random(seed)
    mem = Mem(n)
    adj = [set() for _ in range(n)]
    for a, b in edges:
        adj[a].add(b)
        adj[b].add(a)
    adj = [sorted(s) for s in adj]
    for v in range(n):  # init: p[v] = min(v, smallest neighbour)
        if adj[v] and adj[v][0] < v:
            mem.hist[v] = [adj[v][0]]
    pairs = []
    for v in range(n):  # neighbour 0 when above v, then every other neighbour
        if adj[v] and adj[v][0] > v:
            pairs.append((v, adj[v][0]))
        pairs.extend((v, w) for w in adj[v][1:])
    rng.shuffle(pairs)

    def thread(ps):
        seen = {}
        for a, b in ps:
            yield from rem_unite(mem, seen, a, b, splice, rng, planted)

    gens = [thread(pairs[i::threads]) for i in range(threads)]
    live = list(range(len(gens)))
    steps, cap = 0, 200 * (len(pairs) + n) + 10000
    while live:
        i = rng.choice(live)
        try:
            next(gens[i])
        except StopIteration:
            live.remove(i)
        steps += 1
        assert steps < cap, "step bound exceeded"
    got = []
    for v in range(n):
        x = v
        while mem.latest(x) != x:
            assert mem.latest(x) < x  # p[v] <= v throughout
            x = mem.latest(x)
        got.append(x)
    par = list(range(n))

    def find(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    for a, b in edges:
        ra, rb = find(a), find(b)
        if ra != rb:
            par[max(ra, rb)] = min(ra, rb)
    return got == [find(v) for v in range(n)]


def random_graph(rng):
    n = rng.randint(2, 90)
    edges = [(rng.randrange(n), rng.randrange(n)) for _ in range(rng.randint(0, 3 * n))]
    return n, [(a, b) for a, b in edges if a != b]


@pytest.mark.parametrize("splice", [0, 1, 2])
def test_rem_unite_model(splice):
    for trial in range(250):
        rng = random.Random(1000 * splice + trial)
        n, edges = random_graph(rng)
        assert run(n, edges, rng.randint(1, 40), splice, trial), f"trial {trial}"


def test_rem_unite_model_catches_lost_link():
    failures = 0
    for trial in range(60):
        rng = random.Random(5000 + trial)
        n, edges = random_graph(rng)
        failures += not run(n, edges, rng.randint(2, 40), 0, trial, planted=True)
    assert failures > 0
