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
