"""Parity at the configured full sizes, on the GPU.

* The 65,601,024-vertex Chung-Lu power-law graph (gamma 2.1, 1,921,400,000
  sampled pairs: 1.80e9 undirected edges, 3.60e9 CSR entries after dedup,
  ~15 GB) solved on one GPU, labels bit-exact against the oracle's
  restatement of the reference's label propagation (src/cc_pthreads.c,
  pinned against the reference's own build), with the component count and
  size histogram.
* The 8192 x 8192 lattice with 30% of its edges deleted (high diameter, many
  components), with permuted ids and numbered row by row, against the
  oracle's union-find.
* Rank 0 of a 4-rank vertex-block partition of the same Chung-Lu graph at
  real shape (16,400,256 owned vertices): split, local CC, send lists,
  boundary build and ghosts against their definitions and the oracle, then
  the merge loop with injected ghost deltas at exchange intervals 1, 2 and
  4, every interval-1 pack checked against oracle_delta_pack and the owned
  labels against the full graph's.
"""
import os

import numpy as np
import pytest

import cc_mpi_amd as cm
from cc_mpi_amd import devgen
from cc_mpi_amd.rankshare import RankShare
import oracle

pytestmark = pytest.mark.gpu

CHUNG_LU = dict(n=65601024, pairs=1921400000, gamma=2.1, seed=1003)
THREADS = os.cpu_count() or 1


def histogram(labels):
    sizes = np.bincount(labels, minlength=labels.size)
    sizes = sizes[sizes > 0]
    vals, counts = np.unique(sizes, return_counts=True)
    return int(sizes.size), dict(zip(vals.tolist(), counts.tolist()))


@pytest.fixture(scope="module")
def chung_lu(gpu):
    print("building the Chung-Lu 65.6M graph on the device", flush=True)
    D = devgen.chung_lu(CHUNG_LU["n"], CHUNG_LU["pairs"], CHUNG_LU["gamma"],
                        seed=CHUNG_LU["seed"])
    G = D.to_host()
    print(f"n={G.n} m={G.m}; oracle labels ({THREADS} threads)", flush=True)
    truth = oracle.cc_lp_pthreads(G.row_ptr, G.col_idx, THREADS, 8192)
    yield D, G, truth
    D.close()


def test_chung_lu_65m_full_graph(chung_lu):
    D, G, truth = chung_lu
    assert G.n == 65601024
    assert 3.59e9 < G.m < 3.61e9  # 1.80e9 undirected edges
    got = D.solve()
    bad = np.flatnonzero(got != truth)
    assert bad.size == 0, f"{bad.size} labels differ, first {bad[:5]}"
    count, hist = histogram(got)
    assert (count, hist) == histogram(truth)
    assert count == int(np.count_nonzero(truth == np.arange(G.n, dtype=np.uint32)))
    assert max(hist) > G.n // 2  # a giant component with a tail of small ones
    print(f"components={count} largest={max(hist)} singletons={hist.get(1, 0)}", flush=True)
    # a second solve of the resident graph gives the same labels
    assert np.array_equal(D.solve(), truth)


@pytest.mark.parametrize("permute", [True, False], ids=["permuted", "row_major"])
def test_grid_8192_30pct_deleted(gpu, permute):
    """Both numberings: permuted ids (the sampled path) and row by row (the
    probe picks the vertex-tile plan)."""
    import ctypes
    from cc_mpi_amd import _native as nat
    print(f"building the 8192^2 lattice on the device (permute={permute})", flush=True)
    D = devgen.grid(8192, 8192, 0.7, seed=1004, permute=permute)
    G = D.to_host()
    truth = oracle.cc_min_label(G.row_ptr, G.col_idx)
    got = D.solve()
    ran = ctypes.c_int(-1)
    nat.check(nat.gpu_lib().cc_hip_components_tiles_ran(None, ctypes.byref(ran)), "tiles ran")
    D.close()
    assert ran.value == (0 if permute else 1)
    bad = np.flatnonzero(got != truth)
    assert bad.size == 0, f"{bad.size} labels differ, first {bad[:5]}"
    count, hist = histogram(got)
    assert (count, hist) == histogram(truth)
    assert G.n == 8192 * 8192 and count > 1000  # many components
    print(f"m={G.m} components={count}", flush=True)


@pytest.fixture(scope="module")
def rank0(chung_lu):
    D, G, truth = chung_lu
    rs = RankShare(D, P=4, rank=0)
    rs.preprocess()
    print(f"rank share: {rs.stats} {rs.times}", flush=True)
    yield rs
    rs.close()


def test_rank_share_lists_at_real_shape(chung_lu, rank0):
    """Send lists and ghosts against their definitions, and the sorted
    unique boundary records against the oracle's split, sort, dedup and
    ghost extraction (src/cc_mpi.c:252-461)."""
    D, G, truth = chung_lu
    rs = rank0
    assert (rs.vs, rs.ve) == (0, 16400256)
    nl, blk = rs.nl, 16400256
    lo, hi = int(G.row_ptr[0]), int(G.row_ptr[nl])
    rp = G.row_ptr[: nl + 1]
    ci = G.col_idx[lo:hi]
    deg = np.diff(rp).astype(np.int64)
    src = np.repeat(np.arange(nl, dtype=np.uint32), deg)
    owner = ci // np.uint32(blk)
    # ghosts: every distinct remote column
    mark = np.zeros(G.n, dtype=bool)
    mark[ci[owner != 0]] = True
    ghosts = np.empty(rs.ng, dtype=np.uint32)
    rs.ghosts.download(ghosts)
    assert np.array_equal(ghosts, np.flatnonzero(mark).astype(np.uint32))
    del mark
    # send lists: for each other rank, the owned vertices with a neighbour it owns
    stf = np.empty(rs.total_send, dtype=np.uint32)
    rs.send_to_flat.download(stf)
    assert rs.dests == [1, 2, 3]
    for k, q in enumerate(rs.dests):
        need = np.zeros(nl, dtype=bool)
        need[src[owner == q]] = True
        assert np.array_equal(stf[rs.seg_start[k]:rs.seg_start[k + 1]],
                              np.flatnonzero(need).astype(np.uint32)), f"dest {q}"
    del src, owner
    # boundary records (ghost index << 32 | rep), against the oracle
    comp_of = np.empty(nl, dtype=np.uint32)
    rs.comp_of.download(comp_of)
    rp0 = (rp - rp[0]).astype(np.uint64)
    raw = np.empty(max(1, rs.nb), dtype=np.uint64)
    L = oracle.lib()
    k = L.oracle_boundary_build(nl, oracle._p(rp0, oracle.c_u64), oracle._p(ci, oracle.c_u32),
                                rs.vs, rs.ve, G.n, oracle._p(comp_of, oracle.c_u32),
                                oracle._p(raw, oracle.c_u64))
    assert k == rs.nb
    recs = oracle.dedup_u64(oracle.sort_u64(raw[:k]))
    del raw
    recs, want_ghosts = oracle.ghost_extract(recs)
    assert np.array_equal(want_ghosts, ghosts)
    got = np.empty(rs.nbd, dtype=np.uint64)
    rs.bnd.download(got)
    assert np.array_equal(got, recs)


@pytest.fixture(scope="module")
def truth_on_device(chung_lu):
    buf = cm.DeviceBuffer.from_array(chung_lu[2])
    yield buf
    buf.close()


_runs: dict = {}


@pytest.mark.parametrize("interval", [1, 2, 4])
def test_rank_share_merge_at_real_shape(chung_lu, rank0, truth_on_device, interval):
    D, G, truth = chung_lu
    rs = rank0
    checked = [0]

    def check(stf, seg, cl, co, vs, ve, before, pairs, counts, displs):
        wp, wc, wd = oracle.delta_pack(stf, seg, cl, co, vs, ve, before.copy())
        assert np.array_equal(pairs, wp) and np.array_equal(counts, wc)
        assert np.array_equal(displs, wd)
        checked[0] += 1

    labels = rs.merge(truth_on_device, exchange_interval=interval, seed=11,
                      check_pack=check if interval == 1 else None)
    assert np.array_equal(labels, truth[rs.vs:rs.ve])
    st = dict(rs.stats)
    print(f"interval {interval}: {st}", flush=True)
    slots = rs.prepare_schedule(truth_on_device, seed=11)["last"] + 1
    assert st["exchanges"] >= slots  # every scheduled delta was delivered
    if interval == 1:
        assert checked[0] == st["packs"] >= 2
        assert st["exchanges"] == st["merge_iterations"] - 1
    _runs[interval] = st
    if interval == 4 and {1, 2} <= set(_runs):
        # while the other ranks still send deltas they count as changed, so
        # an exchange waits for `interval` merge rounds (src/cc_mpi.c:669)
        its = [_runs[i]["merge_iterations"] for i in (1, 2, 4)]
        assert its[0] < its[1] < its[2], its
        assert _runs[4]["merge_iterations"] >= 4 * (slots - 1)
