"""Parity of the device boundary-record path (sort, dedup, ghost extraction)
and the delta pack with the oracle, bit-exact."""
import contextlib

import numpy as np
import pytest

import oracle
from cc_mpi_amd import _native as nat
from cc_mpi_amd import boundary as bd
from cc_mpi_amd.cc import DeviceBuffer

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def sort_knobs(**knobs):
    """Set cc_hip_sort_set knobs for the block, restoring the old values."""
    lib = nat.gpu_lib()
    old = {k: lib.cc_hip_sort_get(k.encode()) for k in knobs}
    try:
        for k, v in knobs.items():
            assert lib.cc_hip_sort_set(k.encode(), v) == 0, k
        yield
    finally:
        for k, v in old.items():
            lib.cc_hip_sort_set(k.encode(), v)


def records(count, seed, rep_bits=24, remote_bits=26):
    """Records shaped like one rank's boundary list: rep < 2^rep_bits,
    remote < 2^remote_bits."""
    rng = np.random.default_rng(seed)
    rep = rng.integers(0, 1 << rep_bits, count, dtype=np.uint64)
    remote = rng.integers(0, 1 << remote_bits, count, dtype=np.uint64)
    return (remote << np.uint64(32)) | rep


SIZES = [0, 1, 2, 3, 64, 4095, 4096, 4097, 8191, 8192, 8193, 100_000, 1_000_003]


@pytest.mark.parametrize("count", SIZES)
def test_dropin_sort_matches_oracle(gpu, count):
    k = records(count, seed=count)
    got = k.copy()
    bd.boundary_edges_sort(got)
    assert np.array_equal(got, oracle.sort_u64(k))


# (row-striped tiles) or (a run of items per thread: staged stores, items)
LAYOUTS = [(1, 0, 32), (0, 0, 16), (0, 1, 16), (0, 0, 32)]


@pytest.mark.parametrize("rows,staged,items", LAYOUTS)
@pytest.mark.parametrize("count", SIZES)
def test_dropin_dedup_matches_oracle(gpu, count, rows, staged, items):
    k = oracle.sort_u64(records(count, seed=count + 1, rep_bits=8, remote_bits=9))
    got = k.copy()
    with sort_knobs(compact_rows=rows, dedup_staged=staged, dedup_items=items):
        n = bd.boundary_edges_dedup(got)
    want = oracle.dedup_u64(k)
    assert n == want.size
    assert np.array_equal(got[:n], want)


@pytest.mark.parametrize("kind", ["full64", "equal", "high_only", "low_only", "two_values",
                                  "descending", "sorted"])
def test_sort_key_shapes(gpu, kind):
    rng = np.random.default_rng(7)
    n = 300_001
    if kind == "full64":
        k = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    elif kind == "equal":
        k = np.full(n, 0x0123456789ABCDEF, dtype=np.uint64)
    elif kind == "high_only":
        k = rng.integers(0, 256, n, dtype=np.uint64) << np.uint64(56)
    elif kind == "low_only":
        k = rng.integers(0, 4, n, dtype=np.uint64)
    elif kind == "two_values":
        k = np.where(rng.random(n) < 0.5, np.uint64(5), np.uint64(1 << 40)).astype(np.uint64)
    elif kind == "descending":
        k = np.arange(n, dtype=np.uint64)[::-1].copy() * np.uint64(0x10001)
    else:
        k = np.arange(n, dtype=np.uint64) * np.uint64(977)
    got = k.copy()
    bd.boundary_edges_sort(got)
    assert np.array_equal(got, np.sort(k))


@pytest.mark.parametrize("threads,items", [(256, 8), (256, 16), (256, 24), (256, 32), (512, 8),
                                           (512, 16), (1024, 8)])
@pytest.mark.parametrize("kind", ["records", "skewed", "full64"])
def test_sort_every_tile_size(gpu, threads, items, kind):
    """The scatter passes give the same bytes for every tile shape, on tile
    boundaries and ragged tails, with skewed and uniform digits."""
    lib = nat.gpu_lib()
    rng = np.random.default_rng(items)
    tile = threads * items
    n = tile * 37 + 123
    if kind == "records":
        k = records(n, seed=items)
    elif kind == "skewed":
        k = (rng.integers(0, 3, n, dtype=np.uint64) << np.uint64(56)) | rng.integers(
            0, 1 << 12, n, dtype=np.uint64)
    else:
        k = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    with sort_knobs(items_per_thread=items, sort_threads=threads):
        for m in (n, tile, tile + 1, tile * 5):
            got = k[:m].copy()
            bd.boundary_edges_sort(got)
            assert np.array_equal(got, np.sort(k[:m])), (threads, items, kind, m)


def test_device_sort_ping_pong_and_repeat(gpu):
    k = records(2_000_000, seed=3)
    want = oracle.sort_u64(k)
    a, b = DeviceBuffer.from_array(k), DeviceBuffer(k.nbytes)
    for _ in range(3):
        a.upload(k)
        which = bd.sort_device(a, b, k.size)
        out = np.empty_like(k)
        (b if which else a).download(out)
        assert np.array_equal(out, want)


@pytest.mark.parametrize("rows", [1, 0])
@pytest.mark.parametrize("count", [1_500_000, 4097, 1])
def test_device_dedup_and_ghosts(gpu, rows, count):
    k = oracle.sort_u64(records(count, seed=5, rep_bits=12, remote_bits=14))
    a, b, c = DeviceBuffer.from_array(k), DeviceBuffer(k.nbytes), DeviceBuffer(k.nbytes)
    with sort_knobs(compact_rows=rows):
        n = bd.dedup_device(a, k.size, b)
    uniq = oracle.dedup_u64(k)
    assert n == uniq.size
    got = np.empty(n, dtype=np.uint64)
    b.download(got)
    assert np.array_equal(got, uniq)
    ghosts_d = DeviceBuffer(4 * n)
    with sort_knobs(compact_rows=rows):
        g = bd.ghost_extract_device(b, n, c, ghosts_d)
    want_rec, want_ghosts = oracle.ghost_extract(uniq)
    assert g == want_ghosts.size
    gg = np.empty(g, dtype=np.uint32)
    ghosts_d.download(gg)
    rr = np.empty(n, dtype=np.uint64)
    c.download(rr)
    assert np.array_equal(gg, want_ghosts)
    assert np.array_equal(rr, want_rec)


def test_large_sort_properties(gpu):
    # 200M records of the one-rank boundary shape: sortedness and an
    # order-independent checksum (size-independent properties)
    n = 200_000_000
    rng = np.random.default_rng(11)
    k = (rng.integers(0, 1 << 26, n, dtype=np.uint64) << np.uint64(32)) | \
        rng.integers(0, 1 << 24, n, dtype=np.uint64)
    xs, ss = np.bitwise_xor.reduce(k), int(k.sum(dtype=np.uint64))
    a, b = DeviceBuffer.from_array(k), DeviceBuffer(k.nbytes)
    which = bd.sort_device(a, b, n)
    (b if which else a).download(k)
    assert (k[1:] >= k[:-1]).all()
    assert np.bitwise_xor.reduce(k) == xs and int(k.sum(dtype=np.uint64)) == ss


def delta_case(seed, n_send=50_000, nseg=4, empty=()):
    rng = np.random.default_rng(seed)
    n_local, vs = 40_000, 1_000_000
    ve = vs + n_local
    comp_of = rng.integers(0, n_local, n_local).astype(np.uint32)
    comp_label = rng.integers(0, 2_000_000, n_local).astype(np.uint32)
    stf = rng.integers(vs - 3000, ve + 3000, n_send).astype(np.uint32)
    cuts = np.sort(rng.choice(np.arange(1, n_send), nseg - 1, replace=False)).tolist()
    seg = [0, *cuts, n_send]
    for e in empty:  # make segment e empty
        seg[e + 1] = seg[e]
    seg = np.maximum.accumulate(np.array(seg)).astype(np.int32)
    seg[-1] = n_send
    prev = np.full(n_send, 0xFFFFFFFF, dtype=np.uint32)
    return stf, seg, comp_label, comp_of, vs, ve, prev


# pack designs: three passes (default), or one look-back pass over
# row-striped tiles or runs of items
PACKS = [dict(delta_passes=3), dict(delta_passes=1, compact_rows=1),
         dict(delta_passes=1, compact_rows=0)]


@pytest.mark.parametrize("prepared", [False, True], ids=["gather", "reps"])
@pytest.mark.parametrize("design", PACKS, ids=["three_pass", "lookback_rows", "lookback_runs"])
@pytest.mark.parametrize("seed,nseg,empty", [(0, 1, ()), (1, 4, ()), (2, 5, (0, 2)),
                                             (3, 3, (2,)), (4, 200, ()), (5, 6, (5,))])
def test_delta_pack_rounds(gpu, seed, nseg, empty, design, prepared):
    """Four packs of lowered labels against the oracle's, gathering the reps
    from comp_of each time or reading them from the list gathered once
    (cc_hip_delta_prepare_device); send_to_flat also holds ids outside the
    owned block, whose label is the id itself."""
    stf, seg, cl, co, vs, ve, prev = delta_case(seed, nseg=nseg, empty=empty)
    d_stf, d_cl, d_co = (DeviceBuffer.from_array(x) for x in (stf, cl, co))
    d_prev = DeviceBuffer.from_array(prev)
    sendbuf = bd.PinnedBuffer(stf.size)
    d_rep = None
    if prepared:
        d_rep = DeviceBuffer(4 * stf.size)
        bd.delta_prepare_device(d_stf, stf.size, d_co, vs, ve, d_rep)
        rep = np.empty(stf.size, dtype=np.uint32)
        d_rep.download(rep)
        owned = (stf >= vs) & (stf < ve)
        assert np.array_equal(rep[owned], co[stf[owned] - vs])
        assert (rep[~owned] == 0xFFFFFFFF).all() and (~owned).any()
    rng = np.random.default_rng(seed + 100)
    prev_o = prev.copy()
    for rnd in range(4):
        with sort_knobs(**design):
            pairs, cnt, dis = bd.delta_pack_device(d_stf, stf.size, seg, d_cl, d_co, vs, ve,
                                                   d_prev, sendbuf, d_rep_of_send=d_rep)
        wp, wc, wd = oracle.delta_pack(stf, seg, cl, co, vs, ve, prev_o)
        assert np.array_equal(pairs, wp), f"round {rnd}"
        assert np.array_equal(cnt, wc) and np.array_equal(dis, wd)
        got_prev = np.empty_like(prev)
        d_prev.download(got_prev)
        assert np.array_equal(got_prev, prev_o)
        # labels only go down between exchanges; lower a random subset
        cl = np.minimum(cl, rng.integers(0, 2_000_000, cl.size).astype(np.uint32)
                        | (rng.random(cl.size) > 0.2).astype(np.uint32) * np.uint32(0xFFFFFFFF))
        d_cl.upload(cl)
    sendbuf.close()


@pytest.mark.parametrize("design", PACKS, ids=["three_pass", "lookback_rows", "lookback_runs"])
def test_delta_pack_empty_trailing_segments(gpu, design):
    """Destinations at the end with nothing asked: their segment starts equal
    total_send_to, and their counts and displacements still come out right."""
    stf, seg, cl, co, vs, ve, prev = delta_case(12, nseg=5)
    seg[3:] = stf.size
    d = [DeviceBuffer.from_array(x) for x in (stf, cl, co, prev)]
    sendbuf = bd.PinnedBuffer(stf.size)
    with sort_knobs(**design):
        pairs, cnt, dis = bd.delta_pack_device(d[0], stf.size, seg, d[1], d[2], vs, ve, d[3],
                                               sendbuf)
    wp, wc, wd = oracle.delta_pack(stf, seg, cl, co, vs, ve, prev.copy())
    assert np.array_equal(pairs, wp) and np.array_equal(cnt, wc) and np.array_equal(dis, wd)
    assert cnt[3] == 0 and cnt[4] == 0
    sendbuf.close()


@pytest.mark.parametrize("design", PACKS, ids=["three_pass", "lookback_rows", "lookback_runs"])
def test_delta_pack_nothing_to_send(gpu, design):
    stf, seg, cl, co, vs, ve, prev = delta_case(9)
    prev[:] = 0  # everything already sent at label 0
    d = [DeviceBuffer.from_array(x) for x in (stf, cl, co, prev)]
    sendbuf = bd.PinnedBuffer(stf.size)
    with sort_knobs(**design):
        pairs, cnt, dis = bd.delta_pack_device(d[0], stf.size, seg, d[1], d[2], vs, ve, d[3],
                                               sendbuf)
    assert pairs.size == 0 and cnt.sum() == 0 and (dis == 0).all()


@pytest.mark.parametrize("seed,nsrc", [(0, 1), (1, 3), (2, 7), (3, 64)])
def test_delta_unpack_equals_oracle(gpu, seed, nsrc):
    """cc_hip_delta_unpack_device against oracle_delta_unpack (pinned against
    the reference's src/exchange_plan.c:390-413): received pairs grouped per
    source, some sources sending nothing, positions relative to each
    source's segment of need_from_gidx_flat; over three exchanges, the pairs
    coming from pinned host memory and from device memory."""
    import ctypes
    from cc_mpi_amd import _native as nat
    lib = nat.gpu_lib()
    rng = np.random.default_rng(seed)
    need = rng.integers(0, 3000, nsrc)
    need[rng.random(nsrc) < 0.25] = 0
    tn = int(need.sum())
    ng = tn + 20_000  # ghosts this rank holds; tn of them are needed from the sources
    base = np.concatenate([[0], np.cumsum(need)[:-1]]).astype(np.int32)
    gidx = rng.permutation(ng)[:tn].astype(np.uint32)  # each ghost has one owner
    d_gidx = DeviceBuffer.from_array(gidx if tn else np.zeros(1, np.uint32))
    ghost = rng.integers(0, 1 << 30, ng).astype(np.uint32)
    d_ghost = DeviceBuffer.from_array(ghost)
    want = ghost.copy()
    for rnd in range(3):
        parts, counts = [], []
        for k in range(nsrc):
            take = np.sort(rng.choice(int(need[k]), int(rng.integers(0, need[k] + 1)),
                                      replace=False)) if need[k] else np.zeros(0, np.int64)
            lbl = rng.integers(0, 1 << 30, take.size).astype(np.uint64)
            parts.append((take.astype(np.uint64) << np.uint64(32)) | lbl)
            counts.append(take.size)
        rb = np.concatenate(parts).astype(np.uint64) if parts else np.zeros(0, np.uint64)
        counts = np.array(counts, dtype=np.int32)
        rd = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
        oracle.delta_unpack(rb, counts, rd, base, gidx, want)
        if rnd == 1:  # pairs in device memory
            src = DeviceBuffer.from_array(rb if rb.size else np.zeros(1, np.uint64))
            ptr = src.ptr
        else:  # pairs in pinned host memory, as MPI received them
            src = bd.PinnedBuffer(max(1, rb.size))
            src.array[:rb.size] = rb
            ptr = src.ptr
        nat.check(lib.cc_hip_delta_unpack_device(ptr, rb.size, nat.ptr(rd, ctypes.c_int),
                                                 nat.ptr(base, ctypes.c_int), nsrc, d_gidx.ptr,
                                                 d_ghost.ptr, None), "unpack")
        got = np.empty(ng, dtype=np.uint32)
        d_ghost.download(got)
        assert np.array_equal(got, want), f"exchange {rnd}"
        src.close()
    d_gidx.close()
    d_ghost.close()


@pytest.mark.parametrize("split,rows,items", [(1, 1, 32), (0, 1, 32), (1, 0, 16), (0, 0, 32),
                                               (2, 1, 32), (2, 0, 16)])
@pytest.mark.parametrize("roots", [1, 2, 3, 1000, 65537, "all"])
def test_boundary_build_matches_three_steps(gpu, roots, split, rows, items):
    """cc_hip_boundary_build_device (reps packed by their rank among the
    local roots; with split, the most frequent rep's records through a
    bitmap over the remotes, set by probing (1) or by partitioning the
    remotes into LDS windows (2), merged back) gives the bytes of rep fill
    -> sort -> dedup."""
    import ctypes
    lib = nat.gpu_lib()
    rng = np.random.default_rng(11 if roots == "all" else roots)
    nl = 200_000
    if roots == "all":
        comp_of = np.arange(nl, dtype=np.uint32)
    else:
        rts = np.sort(rng.choice(nl, roots, replace=False)).astype(np.uint32)
        comp_of = rts[rng.integers(0, roots, nl)]
        comp_of[rts] = rts  # every root is its own rep
    count = 3_000_017
    li = rng.integers(0, nl, count, dtype=np.uint64)
    remote = rng.integers(0, 1 << 26, count, dtype=np.uint64)
    rec = (remote << np.uint64(32)) | li
    want = oracle.dedup_u64(oracle.sort_u64((remote << np.uint64(32)) | comp_of[li].astype(np.uint64)))
    d_rec, d_alt = DeviceBuffer.from_array(rec), DeviceBuffer(rec.nbytes)
    d_comp = DeviceBuffer.from_array(comp_of)
    kept, in_alt = ctypes.c_uint64(), ctypes.c_int()
    with sort_knobs(boundary_split=split, compact_rows=rows, dedup_items=items):
        nat.check(lib.cc_hip_boundary_build_device(d_rec.ptr, d_alt.ptr, count, d_comp.ptr, nl,
                                                   ctypes.byref(kept), ctypes.byref(in_alt), None),
                  "bb")
    got = np.empty(int(kept.value), dtype=np.uint64)
    (d_alt if in_alt.value else d_rec).download(got)
    assert np.array_equal(got, want)
    for b in (d_rec, d_alt, d_comp):
        b.close()


@pytest.mark.parametrize("split,split_rows", [(1, 32), (1, 16), (0, 32), (2, 32)])
@pytest.mark.parametrize("count,remote_bits", [(1, 26), (7, 26), (1000, 32), (500_000, 32),
                                               (2_000_003, 26)])
def test_boundary_build_giant_component(gpu, split, split_rows, count, remote_bits):
    """The shape of a real rank's boundary: one local component holds most
    vertices (so most records carry its root), the rest are small; remotes
    up to 2^32 - 1 (with split 2 these need more buckets than the
    partitioned split takes, so it falls back to probing).  Same bytes as
    rep fill -> sort -> dedup."""
    import ctypes
    lib = nat.gpu_lib()
    rng = np.random.default_rng(count + remote_bits)
    nl = 100_000
    comp_of = np.full(nl, 17, dtype=np.uint32)  # the giant component's root is vertex 17
    comp_of[:17] = np.arange(17, dtype=np.uint32)
    small = rng.choice(np.arange(18, nl), 3000, replace=False).astype(np.uint32)
    comp_of[small] = small  # 3000 singleton components
    li = rng.integers(0, nl, count, dtype=np.uint64)
    remote = rng.integers(0, 1 << remote_bits, count, dtype=np.uint64)
    remote[: count // 3] = remote[count // 3: 2 * (count // 3)]  # repeated remotes
    rec = (remote << np.uint64(32)) | li
    want = oracle.dedup_u64(oracle.sort_u64((remote << np.uint64(32)) | comp_of[li].astype(np.uint64)))
    d_rec, d_alt = DeviceBuffer.from_array(rec), DeviceBuffer(rec.nbytes)
    d_comp = DeviceBuffer.from_array(comp_of)
    kept, in_alt = ctypes.c_uint64(), ctypes.c_int()
    with sort_knobs(boundary_split=split, split_dominant_rows=split_rows):
        nat.check(lib.cc_hip_boundary_build_device(d_rec.ptr, d_alt.ptr, count, d_comp.ptr, nl,
                                                   ctypes.byref(kept), ctypes.byref(in_alt), None),
                  "bb")
    got = np.empty(int(kept.value), dtype=np.uint64)
    (d_alt if in_alt.value else d_rec).download(got)
    assert np.array_equal(got, want)
    for b in (d_rec, d_alt, d_comp):
        b.close()
