"""One rank's share of a distributed run on the GPU (split, local CC, send
lists, boundary sort/dedup, ghost extraction, merge with injected deltas,
delta packs): final owned labels equal the global minimum-id labels, every
pack equals the oracle's, and the send/ghost lists equal their definitions."""
import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import devgen
from cc_mpi_amd.rankshare import RankShare, block_bounds
import oracle

pytestmark = pytest.mark.gpu


def graphs():
    return {"chung_lu": lambda: devgen.chung_lu(200000, 900000, 2.1, seed=31),
            "rmat": lambda: devgen.rmat(15, 16, seed=32),
            "grid": lambda: devgen.grid(200, 250, 0.7, seed=33)}


@pytest.fixture(params=[0, 1, 2], ids=["rows", "runs_staged", "runs_direct"])
def store_strategy(request):
    """Every layout of the split and the other compactions: row-striped
    tiles, or a run of items per thread with staged or direct stores."""
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    p = request.param
    knobs = {b"split_staged": int(p == 1), b"dedup_staged": int(p == 1),
             b"compact_rows": int(p == 0)}
    for k, v in knobs.items():
        assert lib.cc_hip_sort_set(k, v) == 0
    yield p
    for k, v in {b"split_staged": 0, b"dedup_staged": 0, b"compact_rows": 1}.items():
        lib.cc_hip_sort_set(k, v)


SPLIT_GRAPHS = {"rmat13": lambda: cm.rmat(13, 16, seed=41),
                "rmat17": lambda: cm.rmat(17, 16, seed=42),
                # long runs of empty rows: most vertices isolated
                "er_sparse": lambda: cm.erdos_renyi(300_000, 0.2, seed=43),
                "grid": lambda: cm.grid(300, 500, 0.7, seed=44)}


@pytest.mark.parametrize("graph,P,rank", [("rmat13", 1, 0), ("rmat13", 4, 0), ("rmat13", 4, 1),
                                          ("rmat13", 7, 6), ("rmat17", 4, 2), ("rmat17", 1, 0),
                                          ("er_sparse", 3, 1), ("grid", 5, 4)])
def test_split_matches_oracle(gpu, store_strategy, graph, P, rank):
    import ctypes
    from cc_mpi_amd import _native as nat
    h = SPLIT_GRAPHS[graph]()
    b = block_bounds(h.n, P)
    vs, ve = int(b[rank]), int(b[rank + 1])
    nl = ve - vs
    # the owned window as the driver sees it (offsets not rebased)
    d_rp = cm.DeviceBuffer.from_array(h.row_ptr)
    d_ci = cm.DeviceBuffer.from_array(h.col_idx)
    lib = nat.gpu_lib()
    mll, nb = ctypes.c_uint64(), ctypes.c_uint64()
    nat.check(lib.cc_hip_split_count_device(d_rp.ptr + 8 * vs, d_ci.ptr, nl, vs, ve, h.n,
                                            ctypes.byref(mll), ctypes.byref(nb), None), "count")
    lrow, boff = cm.DeviceBuffer(8 * (nl + 1)), cm.DeviceBuffer(8 * (nl + 1))
    lcol, rec = cm.DeviceBuffer(max(4, 4 * mll.value)), cm.DeviceBuffer(max(8, 8 * nb.value))
    nat.check(lib.cc_hip_split_emit_device(d_rp.ptr + 8 * vs, d_ci.ptr, nl, vs, ve, h.n, lrow.ptr,
                                           boff.ptr, lcol.ptr, rec.ptr, None), "emit")
    rp = (h.row_ptr[vs:ve + 1] - h.row_ptr[vs]).astype(np.uint64)
    ci = h.col_idx[int(h.row_ptr[vs]):int(h.row_ptr[ve])].copy()
    want_lrow = np.zeros(nl + 1, dtype=np.uint64)
    want_lcol = np.zeros(max(1, ci.size), dtype=np.uint32)
    braw = oracle.c_u64()
    L = oracle.lib()
    m = L.oracle_local_split(nl, oracle._p(rp, oracle.c_u64), oracle._p(ci, oracle.c_u32), vs, ve,
                             h.n, oracle._p(want_lrow, oracle.c_u64),
                             oracle._p(want_lcol, oracle.c_u32), oracle.ctypes.byref(braw))
    assert (mll.value, nb.value) == (m, braw.value)
    got_lrow = np.empty(nl + 1, dtype=np.uint64)
    lrow.download(got_lrow)
    assert np.array_equal(got_lrow, want_lrow)
    got = np.empty(m, dtype=np.uint32)
    lcol.download(got)
    assert np.array_equal(got, want_lcol[:m])
    # boundary records with li (the rep fill maps li -> comp_of[li] later)
    ident = np.arange(nl, dtype=np.uint32)
    want_rec = np.zeros(max(1, ci.size), dtype=np.uint64)
    k = L.oracle_boundary_build(nl, oracle._p(rp, oracle.c_u64), oracle._p(ci, oracle.c_u32), vs,
                                ve, h.n, oracle._p(ident, oracle.c_u32),
                                oracle._p(want_rec, oracle.c_u64))
    got_rec = np.empty(k, dtype=np.uint64)
    rec.download(got_rec)
    assert np.array_equal(got_rec, want_rec[:k])
    got_boff = np.empty(nl + 1, dtype=np.uint64)
    boff.download(got_boff)
    src = np.repeat(np.arange(nl), np.diff(rp).astype(np.int64))
    bmask = ((ci < vs) | (ci >= ve)) & (ci < h.n)
    assert np.array_equal(got_boff, np.concatenate([[0], np.cumsum(np.bincount(
        src[bmask], minlength=nl))]).astype(np.uint64))


@pytest.mark.parametrize("name", ["chung_lu", "rmat", "grid"])
@pytest.mark.parametrize("rank,interval", [(0, 1), (2, 3), (3, 2)])
def test_rank_share(gpu, name, rank, interval):
    full = graphs()[name]()
    h = full.to_host()
    truth = oracle.cc_min_label(h.row_ptr, h.col_idx)
    assert np.array_equal(full.solve(), truth)
    rs = RankShare(full, P=4, rank=rank)
    rs.preprocess()
    vs, ve = rs.vs, rs.ve
    b = block_bounds(h.n, 4)
    assert (vs, ve) == (int(b[rank]), int(b[rank + 1]))

    # send lists and ghosts against their definitions
    rows = np.repeat(np.arange(h.n, dtype=np.int64), np.diff(h.row_ptr).astype(np.int64))
    own = (rows >= vs) & (rows < ve)
    src, dst = rows[own], h.col_idx[own].astype(np.int64)
    remote = (dst < vs) | (dst >= ve)
    owner = np.searchsorted(b, dst[remote], side="right") - 1
    stf = np.empty(rs.total_send, dtype=np.uint32)
    rs.send_to_flat.download(stf)
    for k, q in enumerate(rs.dests):
        want = np.unique(src[remote][owner == q])
        assert np.array_equal(stf[rs.seg_start[k]:rs.seg_start[k + 1]], want), f"dest {q}"
    assert sum(1 for q in range(4) if q != rank and (owner == q).any()) == len(rs.dests)
    gh = np.empty(rs.ng, dtype=np.uint32)
    rs.ghosts.download(gh)
    assert np.array_equal(gh, np.unique(dst[remote]))

    packs = []

    def check(stf, seg, cl, co, vs_, ve_, before, pairs, counts, displs):
        prev = before.copy()
        wp, wc, wd = oracle.delta_pack(stf, seg, cl, co, vs_, ve_, prev)
        assert np.array_equal(pairs, wp) and np.array_equal(counts, wc)
        assert np.array_equal(displs, wd)
        packs.append(pairs.size)

    truth_d = cm.DeviceBuffer.from_array(truth)
    labels = rs.merge(truth_d, exchange_interval=interval, seed=7 + rank, check_pack=check)
    assert np.array_equal(labels, truth[vs:ve])
    assert rs.stats["exchanges"] >= 1 and len(packs) == rs.stats["packs"]
    assert packs[0] == rs.total_send  # the first exchange sends everything
    rs.close()
    full.close()


@pytest.mark.parametrize("rows", [1, 0])
@pytest.mark.parametrize("P,grouped", [(1, True), (2, True), (4, True), (7, False), (64, True),
                                       (64, False)])
def test_send_lists_match_their_definition(gpu, P, grouped, rows):
    """cc_hip_send_lists_device equals, list for list, the sorted unique
    owned ids per destination rank (owner keys -> sort -> dedup ->
    segments), with records grouped by row or in arbitrary order."""
    import ctypes
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    rng = np.random.default_rng(P * 2 + grouped)
    n_global, nl, vs = 5_000_000, 300_000, 1_000_000
    bounds = np.sort(rng.choice(np.arange(1, n_global), P - 1, replace=False)).astype(np.uint32)
    bounds = np.concatenate([[0], bounds, [n_global]]).astype(np.uint32)
    deg = rng.geometric(0.05, nl) - 1
    deg[rng.integers(0, nl, 5)] = 200_000  # a few hub rows spanning many waves
    li = np.repeat(np.arange(nl, dtype=np.uint64), deg)
    remote = rng.integers(0, n_global, li.size, dtype=np.uint64)
    rec = (remote << np.uint64(32)) | li
    if not grouped:
        rec = rec[rng.permutation(rec.size)]
    owner = np.searchsorted(bounds, (rec >> np.uint64(32)).astype(np.int64), side="right") - 1
    d_rec = cm.DeviceBuffer.from_array(rec)
    d_b = cm.DeviceBuffer.from_array(bounds)
    d_vals = cm.DeviceBuffer(4 * max(1, min(rec.size, nl * P)))
    starts = np.zeros(P + 1, dtype=np.int32)
    assert lib.cc_hip_sort_set(b"compact_rows", rows) == 0
    try:
        nat.check(lib.cc_hip_send_lists_device(d_rec.ptr, rec.size, nl, d_b.ptr, P, vs,
                                               nat.ptr(starts, ctypes.c_int), d_vals.ptr, None),
                  "sl")
    finally:
        lib.cc_hip_sort_set(b"compact_rows", 1)
    vals = np.empty(int(starts[P]), dtype=np.uint32)
    d_vals.download(vals)
    low = (rec & np.uint64(0xFFFFFFFF)).astype(np.int64)
    for q in range(P):
        want = np.unique(low[owner == q]) + vs
        assert np.array_equal(vals[starts[q]:starts[q + 1]], want.astype(np.uint32)), q
    # nothing to send
    z = np.zeros(P + 1, dtype=np.int32) - 1
    nat.check(lib.cc_hip_send_lists_device(d_rec.ptr, 0, nl, d_b.ptr, P, vs,
                                           nat.ptr(z, ctypes.c_int), d_vals.ptr, None), "sl0")
    assert not z.any()
    for b in (d_rec, d_b, d_vals):
        b.close()


@pytest.mark.parametrize("nl,roots,nghost,layout",
                         [(1, 1, 3000, "even"), (5000, 37, 3000, "even"), (70_000, 1, 3000, "even"),
                          (5000, 500, 3000, "even"), (20000, 3000, 3000, "even"),
                          (5000, 37, 60000, "even"), (5000, 37, 60000, "clustered"),
                          (5000, 1, 60000, "clustered")])
def test_merge_step_over_rep_list(gpu, nl, roots, nghost, layout):
    """cc_hip_local_reps_device lists exactly {r : comp_of[r] == r} in
    ascending order; merge rounds over that list and rounds that scan
    comp_of both equal the oracle's merge round (src/cc_mpi.c:547-622):
    round for round, with the changed flag, when there is one rep (no other
    rep's update can be seen mid-round), and at convergence otherwise (the
    reps update in parallel, as the reference's threads do, so one round's
    labels depend on the order; the fixed point does not).  With more than
    16K ghosts the hops' ghost lookups bracket the target by interpolation
    first: an even spread of ids takes that path, a clustered one (most ids
    packed at the low end, a sparse tail) falls back to the full search."""
    import ctypes
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    rng = np.random.default_rng(nl + roots + nghost)
    vs, ve = 1000, 1000 + nl
    n_global = ve + max(50_000, 3 * nghost)
    root_ids = np.sort(rng.choice(nl, roots, replace=False)).astype(np.uint32)
    comp_of = root_ids[rng.integers(0, roots, nl)]
    comp_of[root_ids] = root_ids
    if layout == "even":
        ghosts = np.unique(rng.integers(ve, n_global, nghost)).astype(np.uint32)
    else:
        dense = rng.integers(ve, ve + nghost, nghost * 9 // 10)
        tail = rng.integers(ve + nghost, n_global + 40 * nghost, nghost // 10)
        ghosts = np.unique(np.concatenate([dense, tail])).astype(np.uint32)
    # labels: other ghosts' ids half the time, so the hops look ghosts up
    g_lab = np.where(rng.random(ghosts.size) < 0.5, rng.choice(ghosts, ghosts.size),
                     rng.integers(0, n_global, ghosts.size)).astype(np.uint32)
    nb = 20_000
    rec = np.unique((rng.integers(0, ghosts.size, nb).astype(np.uint64) << np.uint64(32))
                    | comp_of[rng.integers(0, nl, nb)].astype(np.uint64))
    d_comp = cm.DeviceBuffer.from_array(comp_of)
    d_reps = cm.DeviceBuffer(4 * nl)
    cnt = ctypes.c_uint32()
    nat.check(lib.cc_hip_local_reps_device(d_comp.ptr, nl, d_reps.ptr, ctypes.byref(cnt), None),
              "reps")
    reps = np.empty(cnt.value, dtype=np.uint32)
    d_reps.download(reps)
    np.testing.assert_array_equal(reps, np.flatnonzero(comp_of == np.arange(nl)).astype(np.uint32))
    d_rec = cm.DeviceBuffer.from_array(rec)
    d_gh = cm.DeviceBuffer.from_array(ghosts)
    prep = _prepare(lib, d_rec, rec, ghosts.size, d_reps, reps)
    d_gmin = cm.DeviceBuffer(4 * reps.size)
    first = [True]

    def gpu_round(use_list, d_gl, d_cl, d_cm, changed):
        if use_list == "prepared":
            rc = lib.cc_hip_merge_step_prepared_device(
                prep["bits"].ptr, prep["dominant"], prep["side"].ptr, prep["n_side"], d_gh.ptr,
                d_gl.ptr, ghosts.size, d_comp.ptr, d_cl.ptr, d_cm.ptr, d_gmin.ptr, nl, d_reps.ptr,
                cnt.value, vs, ve, 2 if first[0] else 0, ctypes.byref(changed), None)
            first[0] = False
        elif use_list:
            rc = lib.cc_hip_merge_step_reps_device(d_rec.ptr, rec.size, d_gh.ptr, d_gl.ptr,
                                                   ghosts.size, d_comp.ptr, d_cl.ptr, d_cm.ptr, nl,
                                                   d_reps.ptr, cnt.value, vs, ve,
                                                   ctypes.byref(changed), None)
        else:
            rc = lib.cc_hip_merge_step_device(d_rec.ptr, rec.size, d_gh.ptr, d_gl.ptr,
                                              ghosts.size, d_comp.ptr, d_cl.ptr, d_cm.ptr, nl, vs,
                                              ve, ctypes.byref(changed), None)
        nat.check(rc, "merge step")

    o_cl = np.arange(vs, ve, dtype=np.uint32)
    o_cm = np.zeros(nl, dtype=np.uint32)
    o_rounds, o_flags = [], []
    while True:
        o_flags.append(oracle.merge_step(rec, ghosts, g_lab, comp_of, o_cl, o_cm, reps, vs, ve))
        o_rounds.append(o_cl[root_ids].copy())
        if not o_flags[-1]:
            break
        assert len(o_rounds) < 100
    for use_list in (False, True, "prepared"):
        first[0] = True
        d_gl = cm.DeviceBuffer.from_array(g_lab)
        d_cl = cm.DeviceBuffer.from_array(np.arange(vs, ve, dtype=np.uint32))
        d_cm = cm.DeviceBuffer(4 * nl)
        changed = ctypes.c_int(-1)
        for rnd in range(100):
            gpu_round(use_list, d_gl, d_cl, d_cm, changed)
            cl = np.empty(nl, dtype=np.uint32)
            d_cl.download(cl)
            if roots == 1:  # one rep: the round is the oracle's, exactly
                np.testing.assert_array_equal(cl[root_ids], o_rounds[min(rnd, len(o_rounds) - 1)],
                                              err_msg=f"round {rnd}")
                assert changed.value == (o_flags[rnd] if rnd < len(o_flags) else 0)
            if not changed.value:
                break
        else:
            raise AssertionError("merge rounds did not converge")
        cmin = np.empty(nl, dtype=np.uint32)
        d_cm.download(cmin)
        # the fixed point: the oracle's, with comp_min equal to comp_label on the reps
        np.testing.assert_array_equal(cl[root_ids], o_cl[root_ids])
        np.testing.assert_array_equal(cmin[root_ids], cl[root_ids])
        # non-reps are never touched
        rest = np.setdiff1d(np.arange(nl), root_ids)
        np.testing.assert_array_equal(cl[rest], np.arange(vs, ve, dtype=np.uint32)[rest])
        for b in (d_gl, d_cl, d_cm):
            b.close()
    # a missing list is an error, not a silent scan
    assert lib.cc_hip_merge_step_reps_device(d_rec.ptr, rec.size, d_gh.ptr, d_gh.ptr, ghosts.size,
                                             d_comp.ptr, d_comp.ptr, d_comp.ptr, nl, None, 0, vs,
                                             ve, ctypes.byref(changed), None) == 4
    for b in (d_comp, d_reps, d_rec, d_gh, d_gmin, prep["bits"], prep["side"]):
        b.close()


def _prepare(lib, d_rec, rec, ghost_count, d_reps, reps):
    """cc_hip_merge_prepare_device over the records, checked against its
    definition: the dominant rep is a rep of the records, its records are
    exactly the set bits, and the other records are the side list in
    input order, each with its rep replaced by the rep's index in the list."""
    import ctypes
    from cc_mpi_amd import _native as nat
    words = (ghost_count + 31) // 32
    d_bits = cm.DeviceBuffer(4 * max(1, words))
    d_side = cm.DeviceBuffer(8 * max(1, rec.size))
    n_side, dom = ctypes.c_uint64(), ctypes.c_uint32()
    nat.check(lib.cc_hip_merge_prepare_device(d_rec.ptr, rec.size, ghost_count, d_reps.ptr,
                                              reps.size, d_bits.ptr, d_side.ptr,
                                              ctypes.byref(n_side), ctypes.byref(dom), None),
              "merge prepare")
    reps_of = (rec & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    gidx = (rec >> np.uint64(32)).astype(np.uint32)
    assert dom.value < reps.size
    is_dom = reps_of == reps[dom.value]
    if rec.size:
        assert is_dom.any()
    bits = np.empty(max(1, words), dtype=np.uint32)
    d_bits.download(bits)
    want_bits = np.zeros(max(1, words), dtype=np.uint32)
    np.bitwise_or.at(want_bits, gidx[is_dom] >> 5, (np.uint32(1) << (gidx[is_dom] & 31)))
    np.testing.assert_array_equal(bits[:words], want_bits[:words])
    side = np.empty(max(1, rec.size), dtype=np.uint64)
    d_side.download(side)
    assert n_side.value == int((~is_dom).sum())
    index_of = np.searchsorted(reps, reps_of[~is_dom]).astype(np.uint64)
    np.testing.assert_array_equal(reps[index_of.astype(np.int64)], reps_of[~is_dom])
    np.testing.assert_array_equal(side[:n_side.value],
                                  (rec[~is_dom] & np.uint64(0xFFFFFFFF00000000)) | index_of)
    return {"bits": d_bits, "side": d_side, "n_side": n_side.value, "dominant": dom.value}


@pytest.mark.parametrize("nl,roots,skew", [(5000, 37, 0.95), (20000, 3000, 0.5), (70_000, 1, 1.0),
                                           (3000, 12, 0.0)])
def test_prepared_rounds_follow_ghost_changes(gpu, nl, roots, skew):
    """Merge rounds over the prepared records keep each rep's minimum ghost
    label until told the ghost labels changed.  A loop that lowers a third
    of the ghost labels after it converges (as an exchange would) and runs
    on to convergence again ends with the labels of the same loop over the
    full record list, and with the oracle's fixed point; with one rep it
    matches the full-list rounds round for round, changed flags included.
    `skew` is the share of records on one rep, as the rank's giant local
    component holds most of them."""
    import ctypes
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    rng = np.random.default_rng(nl * 7 + roots)
    vs, ve, n_global = 2000, 2000 + nl, 2000 + nl + 80_000
    root_ids = np.sort(rng.choice(nl, roots, replace=False)).astype(np.uint32)
    comp_of = root_ids[rng.integers(0, roots, nl)]
    comp_of[root_ids] = root_ids
    ghosts = np.unique(rng.integers(ve, n_global, 6000)).astype(np.uint32)
    g_lab = rng.integers(0, n_global, ghosts.size).astype(np.uint32)
    nb = 40_000
    rep_pick = np.where(rng.random(nb) < skew, root_ids[0], comp_of[rng.integers(0, nl, nb)])
    rec = np.unique((rng.integers(0, ghosts.size, nb).astype(np.uint64) << np.uint64(32))
                    | rep_pick.astype(np.uint64))
    lowered = g_lab.copy()
    pick = rng.random(ghosts.size) < 1 / 3
    lowered[pick] = np.minimum(lowered[pick],
                               rng.integers(0, vs, int(pick.sum())).astype(np.uint32))
    d_comp = cm.DeviceBuffer.from_array(comp_of)
    d_reps = cm.DeviceBuffer(4 * nl)
    cnt = ctypes.c_uint32()
    nat.check(lib.cc_hip_local_reps_device(d_comp.ptr, nl, d_reps.ptr, ctypes.byref(cnt), None),
              "reps")
    d_rec = cm.DeviceBuffer.from_array(rec)
    d_gh = cm.DeviceBuffer.from_array(ghosts)
    reps = np.flatnonzero(comp_of == np.arange(nl)).astype(np.uint32)
    prep = _prepare(lib, d_rec, rec, ghosts.size, d_reps, reps)

    def run(prepared):
        d_gl = cm.DeviceBuffer.from_array(g_lab)
        d_cl = cm.DeviceBuffer.from_array(np.arange(vs, ve, dtype=np.uint32))
        d_cm = cm.DeviceBuffer(4 * nl)
        d_gmin = cm.DeviceBuffer(4 * reps.size)
        changed = ctypes.c_int(-1)
        trace, ghosts_changed = [], 2
        for phase in range(2):
            if phase == 1:
                d_gl.upload(lowered)
                ghosts_changed = 1
            for _ in range(200):
                if prepared:
                    rc = lib.cc_hip_merge_step_prepared_device(
                        prep["bits"].ptr, prep["dominant"], prep["side"].ptr, prep["n_side"],
                        d_gh.ptr, d_gl.ptr, ghosts.size, d_comp.ptr, d_cl.ptr, d_cm.ptr,
                        d_gmin.ptr, nl, d_reps.ptr, cnt.value, vs, ve, ghosts_changed,
                        ctypes.byref(changed), None)
                    ghosts_changed = 0
                else:
                    rc = lib.cc_hip_merge_step_reps_device(
                        d_rec.ptr, rec.size, d_gh.ptr, d_gl.ptr, ghosts.size, d_comp.ptr,
                        d_cl.ptr, d_cm.ptr, nl, d_reps.ptr, cnt.value, vs, ve,
                        ctypes.byref(changed), None)
                nat.check(rc, "merge step")
                cl = np.empty(nl, dtype=np.uint32)
                d_cl.download(cl)
                trace.append((cl[root_ids].copy(), changed.value))
                if not changed.value:
                    break
            else:
                raise AssertionError("merge rounds did not converge")
        for b in (d_gl, d_cl, d_cm, d_gmin):
            b.close()
        return trace

    full, prepd = run(False), run(True)
    np.testing.assert_array_equal(prepd[-1][0], full[-1][0])
    if roots == 1:
        assert len(prepd) == len(full)
        for (a, fa), (b, fb) in zip(full, prepd):
            np.testing.assert_array_equal(a, b)
            assert fa == fb
    # the oracle's fixed point under the lowered labels
    o_cl = np.arange(vs, ve, dtype=np.uint32)
    o_cm = np.zeros(nl, dtype=np.uint32)
    for labels in (g_lab, lowered):
        for _ in range(200):
            if not oracle.merge_step(rec, ghosts, labels, comp_of, o_cl, o_cm, reps, vs, ve):
                break
    np.testing.assert_array_equal(prepd[-1][0], o_cl[root_ids])
    for b in (d_comp, d_reps, d_rec, d_gh, prep["bits"], prep["side"]):
        b.close()


def test_merge_prepare_edge_cases(gpu):
    """No records: the bitmap is cleared (garbage in it beforehand) and the
    side list is empty.  A record whose rep is not in the rep list, and a
    dominant index past the rep list, are errors (code 4), not silent
    out-of-range state."""
    import ctypes
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    reps = np.array([0, 5, 9], dtype=np.uint32)
    d_reps = cm.DeviceBuffer.from_array(reps)
    d_bits = cm.DeviceBuffer.from_array(np.full(4, 0xDEADBEEF, dtype=np.uint32))
    d_side = cm.DeviceBuffer(8 * 8)
    n_side, dom = ctypes.c_uint64(7), ctypes.c_uint32(7)
    nat.check(lib.cc_hip_merge_prepare_device(None, 0, 100, d_reps.ptr, reps.size, d_bits.ptr,
                                              d_side.ptr, ctypes.byref(n_side), ctypes.byref(dom),
                                              None), "prepare, no records")
    bits = np.empty(4, dtype=np.uint32)
    d_bits.download(bits)
    assert n_side.value == 0 and dom.value == 0 and not bits.any()
    rec = np.array([(1 << 32) | 5, (2 << 32) | 5, (3 << 32) | 6], dtype=np.uint64)  # 6 is no rep
    d_rec = cm.DeviceBuffer.from_array(rec)
    rc = lib.cc_hip_merge_prepare_device(d_rec.ptr, rec.size, 100, d_reps.ptr, reps.size,
                                         d_bits.ptr, d_side.ptr, ctypes.byref(n_side),
                                         ctypes.byref(dom), None)
    assert rc == 4 and b"not in the rep list" in lib.cc_hip_last_error()
    d_one = cm.DeviceBuffer.from_array(np.zeros(16, dtype=np.uint32))
    changed = ctypes.c_int()
    rc = lib.cc_hip_merge_step_prepared_device(
        d_bits.ptr, 3, d_side.ptr, 0, d_one.ptr, d_one.ptr, 1, d_one.ptr, d_one.ptr, d_one.ptr,
        d_one.ptr, 10, d_reps.ptr, reps.size, 0, 10, 2, ctypes.byref(changed), None)
    assert rc == 4 and b"dominant index" in lib.cc_hip_last_error()
    for b in (d_reps, d_bits, d_side, d_rec, d_one):
        b.close()
