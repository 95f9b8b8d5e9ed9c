"""Model of the vertex-tile plan (k_tile_local, k_tile_cross, k_tile_roots
and the last pass in components.hip), run on the CPU with the device's data
layout: tiles of 2^shift consecutive ids, a pair segment of `cap` entries
per tile whose count may exceed the cap (then the cross pass reads the
tile's rows again), a local-root bitmap, and a direct-mapped table of the
pairs each tile's cross pass has already united.

Concurrency is modelled where the design depends on it:
  * the cross passes of all tiles interleave pair by pair, each pair read
    on the latest parents;
  * in the roots pass, other walks' halving stores land on a local root's
    entry after its own final store; they are atomicMin, so they cannot
    raise it and every local root ends holding exactly its root, which is
    what lets the last pass take a single hop.
Labels must be each component's minimum.  Two planted mistakes must fail:
plain halving stores in the roots pass (one may leave a local root an
ancestor short of its root), and a cross pass that ignores a tile whose
pairs overflowed its segment.
"""
import random

import pytest


def find(p, x):
    while p[x] != x:
        assert p[x] < x
        x = p[x]
    return x


def hook(p, a, b):
    ra, rb = find(p, a), find(p, b)
    if ra != rb:
        p[max(ra, rb)] = min(ra, rb)


def tile_solve(n, rows, shift, cap, rng, sym=True, plain_stores=False, drop_overflow=False):
    T = 1 << shift
    nt = (n + T - 1) // T
    p = list(range(n))
    bits = [0] * ((n + 31) // 32)
    counts = [0] * nt
    segs = [[] for _ in range(nt)]
    order = list(range(nt))
    rng.shuffle(order)
    for t in order:  # k_tile_local: tiles are independent
        t0, tn = t * T, min(n - t * T, T)
        par = list(range(tn))
        entries = [(vl, w) for vl in range(tn) for w in rows[t0 + vl]]
        rng.shuffle(entries)  # the waves' unions land in any order
        for vl, w in entries:
            if w >= n:
                continue
            wl = w - t0
            if 0 <= wl < tn:
                if wl != vl and (not sym or wl < vl):
                    hook(par, vl, wl)
            elif not sym or w < t0:
                if counts[t] < cap:
                    segs[t].append((t0 + vl, w))
                counts[t] += 1
        for i in range(tn):
            r = find(par, i)
            p[t0 + i] = t0 + r
            if r == i:
                bits[(t0 + i) >> 5] |= 1 << ((t0 + i) & 31)

    # k_tile_cross: every tile's work as a queue of pairs, interleaved
    work = {}
    for t in range(nt):
        t0, tn = t * T, min(n - t * T, T)
        if counts[t] <= cap:
            work[t] = list(segs[t])
        elif drop_overflow:
            work[t] = list(segs[t])  # planted mistake: only the pairs that fit
        else:
            work[t] = [(t0 + vl, w) for vl in range(tn) for w in rows[t0 + vl]
                       if w < n and not (0 <= w - t0 < tn) and (not sym or w < t0)]
    seen = {t: {} for t in range(nt)}
    live = [t for t in range(nt) if work[t]]
    while live:
        t = rng.choice(live)
        v, w = work[t].pop()
        if not work[t]:
            live.remove(t)
        a, b = p[v], p[w]
        if a == b:
            continue
        h = hash((a, b)) % 16  # a small table, so that slots are overwritten
        if seen[t].get(h) == (a, b):
            continue
        seen[t][h] = (a, b)
        hook(p, a, b)

    # k_tile_roots: each linked local root's walk; then halving stores of
    # other walks through it land after its own final store, each carrying
    # an ancestor it had on the forest this pass started from
    roots = [i * 32 + b for i, word in enumerate(bits) for b in range(32) if word >> b & 1]
    rng.shuffle(roots)
    paths = {}
    for r in roots:
        path = [r]
        while p[path[-1]] != path[-1]:
            path.append(p[path[-1]])
        paths[r] = path
    for r in roots:
        if len(paths[r]) > 1:
            p[r] = find(p, r)
    for r in roots:
        if len(paths[r]) > 2 and rng.random() < 0.5:
            late = rng.choice(paths[r][1:])
            p[r] = late if plain_stores else min(p[r], late)

    # the last pass, one hop, in no particular order (in increasing order
    # every local root, the smallest id of its tile component, would be done
    # before its members read it)
    vs = list(range(n))
    rng.shuffle(vs)
    for v in vs:
        x = p[v]
        if x != v:
            p[v] = p[x]
    return p


def components(n, rows):
    par = list(range(n))

    def f(x):
        while par[x] != x:
            par[x] = par[par[x]]
            x = par[x]
        return x

    for v in range(n):
        for w in rows[v]:
            if w < n:
                a, b = f(v), f(w)
                if a != b:
                    par[max(a, b)] = min(a, b)
    return [f(v) for v in range(n)]


def lattice(rng, r, c, keep, permute=False):
    n = r * c
    ids = list(range(n))
    if permute:
        rng.shuffle(ids)
    rows = [[] for _ in range(n)]
    for i in range(r):
        for j in range(c):
            v = i * c + j
            for w in ((v + 1) if j + 1 < c else None, (v + c) if i + 1 < r else None):
                if w is not None and rng.random() < keep:
                    rows[ids[v]].append(ids[w])
                    rows[ids[w]].append(ids[v])
    for row in rows:
        row.sort()
    return n, rows


def random_graph(rng):
    n = rng.randint(1, 300)
    rows = [[] for _ in range(n)]
    for _ in range(rng.randint(0, 3 * n)):
        a = rng.randrange(n)
        b = (a + rng.randint(-40, 40)) % n if rng.random() < 0.8 else rng.randrange(n)
        if a != b:
            rows[a].append(b)
            rows[b].append(a)
    if rng.random() < 0.3 and n > 1:  # duplicates and entries past n
        rows[0].append(n + 5)
        rows[n - 1].append(0)
        rows[0].append(n - 1)
    for row in rows:
        rng.shuffle(row)
    return n, rows


@pytest.mark.parametrize("shift", [6, 7, 9])
def test_tile_model_gives_component_minima(shift):
    for trial in range(120):
        rng = random.Random(trial * 10 + shift)
        if trial % 3 == 0:
            n, rows = lattice(rng, rng.randint(1, 30), rng.randint(1, 40), rng.choice([0.3, 0.7, 1.0]),
                              permute=trial % 2 == 1)
        else:
            n, rows = random_graph(rng)
        cap = rng.choice([1, 4, 1 << shift])
        got = tile_solve(n, rows, shift, cap, rng)
        assert got == components(n, rows), f"trial {trial}"


def test_tile_model_one_directional():
    """symmetric = 0: every stored entry is united (in the tile, or sent as a
    pair whatever its direction), giving the weakly connected components."""
    for trial in range(100):
        rng = random.Random(500 + trial)
        n, rows = random_graph(rng)
        for row in rows:  # drop one direction of some entries
            for k in range(len(row) - 1, -1, -1):
                if rng.random() < 0.3:
                    del row[k]
        got = tile_solve(n, rows, 6, rng.choice([1, 64]), rng, sym=False)
        assert got == components(n, rows), f"trial {trial}"


def test_tile_model_catches_planted_mistakes():
    plain = overflow = 0
    for trial in range(150):
        rng = random.Random(900 + trial)
        n, rows = lattice(rng, rng.randint(4, 30), rng.randint(8, 40), 0.8)
        want = components(n, rows)
        plain += tile_solve(n, rows, 6, 64, random.Random(trial), plain_stores=True) != want
        overflow += tile_solve(n, rows, 6, 2, random.Random(trial), drop_overflow=True) != want
    assert plain > 0 and overflow > 0
