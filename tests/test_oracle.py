"""Pins the oracle (CPU restatement) before it is trusted as the checker.

Each oracle function is checked against something independent of its own
formula: the reference itself compiled from its sources (oracle/_ref), a
brute-force search, numpy/scipy library routines, or closed-form facts.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.sparse.csgraph import connected_components

import cc_mpi_amd as cm
import oracle
from graphs_util import bfs_min_labels, complete, csr_raw, disjoint_cliques, path, star

SMALL = [
    lambda: path(50),
    lambda: path(300, perm_seed=3),
    lambda: star(40, centre=17),
    lambda: complete(12),
    lambda: disjoint_cliques(9, 5, seed=2),
    lambda: cm.erdos_renyi(500, 1.2, seed=4),
    lambda: cm.grid(20, 20, 0.55, seed=8),
    lambda: cm.rmat(9, 4, seed=1),
]


def scipy_min_labels(G):
    n = G.n
    A = sp.csr_matrix((np.ones(G.m, dtype=np.int8), G.col_idx.astype(np.int64),
                       G.row_ptr.astype(np.int64)), shape=(n, n))
    k, comp = connected_components(A, directed=False)
    mins = np.full(k, n, dtype=np.int64)
    np.minimum.at(mins, comp, np.arange(n))
    return mins[comp].astype(np.uint32)


@pytest.mark.parametrize("make", SMALL)
def test_min_label_vs_bfs(make):
    G = make()
    assert np.array_equal(oracle.cc_min_label(G.row_ptr, G.col_idx), bfs_min_labels(G))


@pytest.mark.parametrize("seed", [1, 2])
def test_min_label_vs_scipy(seed):
    for G in (cm.rmat(14, 8, seed=seed), cm.erdos_renyi(20000, 1.5, seed=seed),
              cm.grid(100, 120, 0.6, seed=seed)):
        assert np.array_equal(oracle.cc_min_label(G.row_ptr, G.col_idx), scipy_min_labels(G))


def test_min_label_raw_csr_edge_cases():
    # duplicates, self-loops, out-of-range entries (ignored), isolated vertex
    G = csr_raw(6, [[1, 1, 0], [0, 0], [9], [4], [3, 3], []])
    assert oracle.cc_min_label(G.row_ptr, G.col_idx).tolist() == [0, 0, 2, 3, 3, 5]
    G = csr_raw(1, [[]])
    assert oracle.cc_min_label(G.row_ptr, G.col_idx).tolist() == [0]


def test_min_label_closed_form_properties():
    G = cm.rmat(13, 8, seed=9)
    lab = oracle.cc_min_label(G.row_ptr, G.col_idx)
    v = np.arange(G.n, dtype=np.uint32)
    assert (lab <= v).all()
    assert (lab[lab] == lab).all()
    src = np.repeat(v, np.diff(G.row_ptr).astype(np.int64))
    assert (lab[src] == lab[G.col_idx]).all()


needs_ref = pytest.mark.skipif(oracle.ref_lib() is None,
                               reason="oracle/_ref not built (reference sources absent)")


@needs_ref
@pytest.mark.parametrize("threads,chunk", [(1, 4096), (4, 1), (4, 8192), (3, 0), (8, 2048)])
def test_lp_restatement_matches_reference(threads, chunk):
    for G in (cm.erdos_renyi(65536, 8.0, seed=1001), cm.rmat(14, 16, seed=3),
              cm.grid(64, 64, 0.7, seed=5)):
        ref = oracle.ref_cc_pthreads(G.row_ptr, G.col_idx, threads, chunk)
        mine = oracle.cc_lp_pthreads(G.row_ptr, G.col_idx, threads, chunk)
        uf = oracle.cc_min_label(G.row_ptr, G.col_idx)
        assert np.array_equal(ref, uf)
        assert np.array_equal(mine, uf)


def test_reference_needs_both_directions():
    """The reference's propagation (src/cc_pthreads.c:32-73) pushes a label
    to the neighbours only when the vertex itself just lowered it, so on a
    CSR that stores some edges in one direction only its fixed point depends
    on the schedule: thread count and chunk size change the labels, and none
    of them is the weakly connected components.  With both directions stored
    every schedule gives the per-component minimum.  This is why the drop-in
    takes symmetric input as the contract (the knob symmetric = 0 gives the
    weak components instead, which no reference schedule does)."""
    from graphs_util import one_directional
    if oracle.ref_lib() is None:
        pytest.skip("reference build absent")
    schedules = ((1, 4096), (4, 1), (8, 64), (8, 4096), (3, 7))
    varied = 0
    for seed, G in enumerate((cm.rmat(12, 8, seed=21), cm.grid(48, 48, 0.7, seed=22),
                              cm.erdos_renyi(20000, 2.5, seed=23))):
        uf = oracle.cc_min_label(G.row_ptr, G.col_idx)
        for t, c in schedules:
            assert np.array_equal(oracle.ref_cc_pthreads(G.row_ptr, G.col_idx, t, c), uf)
        D = one_directional(G, 0.2, seed)
        assert D.m < G.m
        weak = oracle.cc_min_label(D.row_ptr, D.col_idx)
        assert np.array_equal(weak, uf)  # dropping a direction keeps weak connectivity
        outs = [oracle.ref_cc_pthreads(D.row_ptr, D.col_idx, t, c) for t, c in schedules]
        assert not any(np.array_equal(o, weak) for o in outs)
        varied += any(not np.array_equal(o, outs[0]) for o in outs[1:])
    assert varied >= 1


def test_lp_restatement_without_reference():
    G = cm.rmat(12, 16, seed=2)
    for t, c in ((1, 4096), (5, 1), (6, 7)):
        assert np.array_equal(oracle.cc_lp_pthreads(G.row_ptr, G.col_idx, t, c),
                              oracle.cc_min_label(G.row_ptr, G.col_idx))


def random_records(count, seed, rep_bits=20, remote_bits=26):
    rng = np.random.default_rng(seed)
    rep = rng.integers(0, 1 << rep_bits, count, dtype=np.uint64)
    remote = rng.integers(0, 1 << remote_bits, count, dtype=np.uint64)
    return (remote << np.uint64(32)) | rep


@pytest.mark.parametrize("count", [0, 1, 2, 63, 64, 65, 1000, 200000])
def test_sort_vs_numpy(count):
    k = random_records(count, seed=count)
    assert np.array_equal(oracle.sort_u64(k), np.sort(k))


def test_sort_full_width_and_constant_digits():
    rng = np.random.default_rng(5)
    k = rng.integers(0, np.iinfo(np.uint64).max, 5000, dtype=np.uint64, endpoint=True)
    assert np.array_equal(oracle.sort_u64(k), np.sort(k))
    k = np.full(100, 0xABCD00000000EF01, dtype=np.uint64)
    assert np.array_equal(oracle.sort_u64(k), k)


@needs_ref
@pytest.mark.parametrize("count", [2, 64, 65, 5000, 300000])
def test_sort_and_dedup_match_reference(count):
    k = random_records(count, seed=11 + count, rep_bits=6, remote_bits=8)  # many duplicates
    ref = k.copy()
    L = oracle.ref_lib()
    L.boundary_edges_sort(oracle._p(ref, oracle.c_u64), ref.size)
    assert np.array_equal(oracle.sort_u64(k), ref)
    nd = L.boundary_edges_dedup(oracle._p(ref, oracle.c_u64), ref.size)
    mine = oracle.dedup_u64(oracle.sort_u64(k))
    assert np.array_equal(mine, ref[:nd])
    assert np.array_equal(mine, np.unique(k))


def test_ghost_extract_vs_numpy():
    k = np.unique(random_records(20000, seed=7, rep_bits=10, remote_bits=12))
    rec, ghosts = oracle.ghost_extract(k)
    remote = (k >> np.uint64(32)).astype(np.uint32)
    uq, inv = np.unique(remote, return_inverse=True)
    assert np.array_equal(ghosts, uq)
    assert np.array_equal((rec >> np.uint64(32)).astype(np.int64), inv)
    assert np.array_equal(rec & np.uint64(0xFFFFFFFF), k & np.uint64(0xFFFFFFFF))


def brute_delta(stf, seg, comp_label, comp_of, vs, ve, prev):
    pairs, counts = [], []
    for k in range(len(seg) - 1):
        c = 0
        for i in range(seg[k + 1] - seg[k]):
            idx = seg[k] + i
            v = int(stf[idx])
            lbl = int(comp_label[comp_of[v - vs]]) if vs <= v < ve else v
            if lbl < prev[idx]:
                prev[idx] = lbl
                pairs.append((i << 32) | lbl)
                c += 1
        counts.append(c)
    return np.array(pairs, dtype=np.uint64), np.array(counts, dtype=np.int32)


@pytest.mark.parametrize("seed", range(4))
def test_delta_pack_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n_local, vs = 300, 1000
    ve = vs + n_local
    comp_of = rng.integers(0, n_local, n_local).astype(np.uint32)
    comp_label = rng.integers(0, 5000, n_local).astype(np.uint32)
    stf = rng.integers(vs - 50, ve + 50, 700).astype(np.uint32)  # some not owned
    cuts = np.sort(rng.choice(np.arange(1, 700), 3, replace=False))
    seg = [0, *cuts.tolist(), 700]
    prev = np.full(700, 0xFFFFFFFF, dtype=np.uint32)
    prev[rng.random(700) < 0.3] = rng.integers(0, 3000, 1).astype(np.uint32)[0]
    prev_b = prev.astype(np.int64).tolist()
    pairs, cnt, dis = oracle.delta_pack(stf, seg, comp_label, comp_of, vs, ve, prev)
    want_pairs, want_cnt = brute_delta(stf, seg, comp_label, comp_of, vs, ve, prev_b)
    assert np.array_equal(pairs, want_pairs)
    assert np.array_equal(cnt, want_cnt)
    assert np.array_equal(dis, np.concatenate([[0], np.cumsum(want_cnt)[:-1]]).astype(np.int32))
    assert np.array_equal(prev.astype(np.int64), np.array(prev_b))
    # a second pack with nothing changed emits nothing
    pairs2, cnt2, _ = oracle.delta_pack(stf, seg, comp_label, comp_of, vs, ve, prev)
    assert pairs2.size == 0 and cnt2.sum() == 0


def test_partition_math():
    L = oracle.lib()
    for n in (1, 7, 100, 1001):
        for P in (1, 2, 3, 4, 8):
            if P > n:
                continue
            starts, ends = [], []
            for r in range(P):
                a, b = oracle.c_u32(), oracle.c_u32()
                L.oracle_vertex_range(n, P, r, oracle.ctypes.byref(a), oracle.ctypes.byref(b))
                starts.append(a.value)
                ends.append(b.value)
            assert starts[0] == 0 and ends[-1] == n
            assert all(ends[i] == starts[i + 1] for i in range(P - 1))
            sizes = [e - s for s, e in zip(starts, ends)]
            assert max(sizes) - min(sizes) <= 1 and sizes == sorted(sizes, reverse=True)
            for v in range(n):
                own = L.oracle_owner_of_vertex(v, n, P, None, 0)
                assert starts[own] <= v < ends[own]
    G = cm.rmat(10, 8, seed=3)
    for P in (2, 3, 4, 7):
        b = np.zeros(P + 1, dtype=np.uint32)
        L.oracle_weighted_bounds(oracle._p(G.row_ptr, oracle.c_u64), G.n, P, 1,
                                 oracle._p(b, oracle.c_u32))
        assert b[0] == 0 and b[-1] == G.n and (np.diff(b.astype(np.int64)) >= 1).all()
        w = np.diff(G.row_ptr).astype(np.int64) + 1
        cum = np.concatenate([[0], np.cumsum(w)])
        for r in range(1, P):
            # first vertex whose prefix weight reaches r/P of the total
            target = cum[-1] * r // P
            assert b[r] == max(int(np.searchsorted(cum, target, side="left")), int(b[r - 1]) + 1)
        for v in range(G.n):
            own = L.oracle_owner_of_vertex(v, G.n, P, oracle._p(b, oracle.c_u32), 1)
            assert b[own] <= v < b[own + 1]


def test_local_split_and_boundary_build():
    G = cm.rmat(11, 8, seed=6)
    n, vs, ve = G.n, 500, 1300
    rows = slice(vs, ve + 1)
    rp = (G.row_ptr[rows] - G.row_ptr[vs]).astype(np.uint64)
    ci = G.col_idx[int(G.row_ptr[vs]):int(G.row_ptr[ve])].copy()
    nl = ve - vs
    lrow = np.zeros(nl + 1, dtype=np.uint64)
    lcol = np.zeros(max(1, ci.size), dtype=np.uint32)
    braw = oracle.c_u64()
    L = oracle.lib()
    mll = L.oracle_local_split(nl, oracle._p(rp, oracle.c_u64), oracle._p(ci, oracle.c_u32), vs,
                               ve, n, oracle._p(lrow, oracle.c_u64), oracle._p(lcol, oracle.c_u32),
                               oracle.ctypes.byref(braw))
    src = np.repeat(np.arange(nl), np.diff(rp).astype(np.int64))
    loc = (ci >= vs) & (ci < ve)
    assert mll == loc.sum() and braw.value == (~loc).sum()
    assert np.array_equal(lcol[:mll], ci[loc] - vs)
    assert np.array_equal(np.diff(lrow).astype(np.int64), np.bincount(src[loc], minlength=nl))
    comp_of = np.random.default_rng(0).integers(0, nl, nl).astype(np.uint32)
    rec = np.zeros(max(1, ci.size), dtype=np.uint64)
    nb = L.oracle_boundary_build(nl, oracle._p(rp, oracle.c_u64), oracle._p(ci, oracle.c_u32), vs,
                                 ve, n, oracle._p(comp_of, oracle.c_u32),
                                 oracle._p(rec, oracle.c_u64))
    want = (ci[~loc].astype(np.uint64) << np.uint64(32)) | comp_of[src[~loc]].astype(np.uint64)
    assert np.array_equal(rec[:nb], want)


# ---- merge round (src/cc_mpi.c:547-622) ------------------------------------

def _rec(gidx, rep):
    return (np.uint64(gidx) << np.uint64(32)) | np.uint64(rep)


def test_merge_step_hand_cases():
    """Worked by hand from src/cc_mpi.c:547-622.  Block [50, 60), reps 0 and
    5; ghosts 30, 40 with labels 40, 52.  Rep 0 takes 40 from its record
    and stops (label of 40 is 52).  Rep 5, after it, takes 52 from its
    record, hops to label_of(52) = comp_label[comp_of[2]] = 40 (already
    lowered this round), then stops.  In the other list order rep 5 first
    hops to 50 only."""
    vs, ve = 50, 60
    comp_of = np.array([0] * 5 + [5] * 5, dtype=np.uint32)
    gv = np.array([30, 40], dtype=np.uint32)
    gl = np.array([40, 52], dtype=np.uint32)
    rec = np.array([_rec(0, 0), _rec(1, 5)], dtype=np.uint64)
    for order, want in (([0, 5], (40, 40)), ([5, 0], (40, 50))):
        cl = np.arange(vs, ve, dtype=np.uint32)
        cm = np.zeros(10, dtype=np.uint32)
        assert oracle.merge_step(rec, gv, gl, comp_of, cl, cm, np.array(order), vs, ve) == 1
        assert (int(cl[0]), int(cl[5])) == want
        assert oracle.merge_step(rec, gv, gl, comp_of, cl, cm, np.array(order), vs, ve) == \
            int(want[1] != 40)
        assert (int(cl[0]), int(cl[5])) == (40, 40)  # the next round agrees
    # four hops at most: rep 0's record gives 44 (ghost 45's label); the
    # ghost chain 44 -> 43 -> 42 -> 41 -> 40 is four hops, and with one more
    # link (40 -> 39) the fifth hop is not taken
    for gv, gl, want in (([40, 41, 42, 43, 44, 45], [40, 40, 41, 42, 43, 44], 40),
                         ([39, 40, 41, 42, 43, 44, 45], [39, 39, 40, 41, 42, 43, 44], 40)):
        gv, gl = np.array(gv, dtype=np.uint32), np.array(gl, dtype=np.uint32)
        rec = np.array([_rec(gv.size - 1, 0)], dtype=np.uint64)
        cl = np.arange(vs, ve, dtype=np.uint32)
        cm = np.zeros(10, dtype=np.uint32)
        oracle.merge_step(rec, gv, gl, comp_of, cl, cm, np.array([0, 5]), vs, ve)
        assert int(cl[0]) == want
    # a vertex that is neither owned nor a ghost ends the hops ("missing")
    gv3 = np.array([45], dtype=np.uint32)
    gl3 = np.array([44], dtype=np.uint32)
    rec = np.array([_rec(0, 5)], dtype=np.uint64)
    cl = np.arange(vs, ve, dtype=np.uint32)
    assert oracle.merge_step(rec, gv3, gl3, comp_of, cl, cm, np.array([0, 5]), vs, ve) == 1
    assert int(cl[5]) == 44 and int(cl[0]) == 50


def _share(G, P, rank):
    """Rank `rank` of a P-way block partition: comp_of (local roots),
    boundary records (gidx << 32) | rep sorted unique, ghosts, reps."""
    n = G.n
    base, rem = divmod(n, P)
    sizes = [base + 1 if r < rem else base for r in range(P)]
    b = np.concatenate([[0], np.cumsum(sizes)])
    vs, ve = int(b[rank]), int(b[rank + 1])
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(G.row_ptr).astype(np.int64))
    own = (rows >= vs) & (rows < ve)
    src, dst = rows[own] - vs, G.col_idx[own].astype(np.int64)
    loc = (dst >= vs) & (dst < ve)
    nl = ve - vs
    lr = np.concatenate([[0], np.cumsum(np.bincount(src[loc], minlength=nl))]).astype(np.uint64)
    lc = (dst[loc] - vs).astype(np.uint32)
    comp_of = oracle.cc_min_label(lr, lc)
    ghosts, gidx = np.unique(dst[~loc], return_inverse=True)
    rec = np.unique((gidx.astype(np.uint64) << np.uint64(32)) | comp_of[src[~loc]].astype(np.uint64))
    reps = np.flatnonzero(comp_of == np.arange(nl)).astype(np.uint32)
    return vs, ve, comp_of, rec, ghosts.astype(np.uint32), reps


@pytest.mark.parametrize("P,rank", [(2, 0), (4, 1), (4, 3), (7, 5)])
def test_merge_step_closed_form(P, rank):
    """With every ghost holding its true global label (the component
    minimum), one merge round gives each owned vertex its global label
    comp_label[comp_of[li]] = min id of its global component."""
    import cc_mpi_amd as cm
    G = cm.rmat(12, 4, seed=P * 10 + rank)
    truth = oracle.cc_min_label(G.row_ptr, G.col_idx)
    vs, ve, comp_of, rec, ghosts, reps = _share(G, P, rank)
    cl = np.arange(vs, ve, dtype=np.uint32)
    cmin = np.zeros(ve - vs, dtype=np.uint32)
    oracle.merge_step(rec, ghosts, truth[ghosts], comp_of, cl, cmin, reps, vs, ve)
    assert np.array_equal(cl[comp_of], truth[vs:ve])
    assert oracle.merge_step(rec, ghosts, truth[ghosts], comp_of, cl, cmin, reps, vs, ve) == 0


def test_merge_step_converges_whatever_the_order():
    """Rounds repeated until nothing changes reach the same labels with the
    rep list forwards and backwards, from intermediate ghost labels (each
    ghost's label on its own rank after local CC): the order only changes
    how many rounds it takes."""
    import cc_mpi_amd as cm
    for seed in range(6):
        G = cm.rmat(11, 3, seed=100 + seed)
        P, rank = 4, seed % 4
        vs, ve, comp_of, rec, ghosts, reps = _share(G, P, rank)
        # a ghost's intermediate label: the smallest id it reaches within its
        # own block (what its owner sends first)
        inter = ghosts.copy()
        for q in range(P):
            qs, qe, qco, _, _, _ = _share(G, P, q)
            m = (ghosts >= qs) & (ghosts < qe)
            inter[m] = qco[ghosts[m] - qs] + qs
        ends = []
        for order in (reps, reps[::-1].copy()):
            cl = np.arange(vs, ve, dtype=np.uint32)
            cmin = np.zeros(ve - vs, dtype=np.uint32)
            rounds = 0
            while oracle.merge_step(rec, ghosts, inter, comp_of, cl, cmin, order, vs, ve):
                rounds += 1
                assert rounds < 100
            ends.append(cl[comp_of].copy())
        assert np.array_equal(ends[0], ends[1]), seed
