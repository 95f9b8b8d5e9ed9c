"""One rank's share of a distributed CC run, at its real shape, on one GPU.

The per-rank path of the reference's driver (src/cc_mpi.c:221-763) for rank
`rank` of a P-rank vertex-block partition of a graph resident in HBM, with
the MPI transport replaced by a seeded injection schedule: every exchange
interval the ghosts receive label updates (each ghost reaches its true
global label at a seeded exchange index, with an intermediate value half
way), exactly as deltas from the other ranks would arrive.  Everything the
rank computes runs on the GPU through the C ABI:

  split (local CSR + raw boundary records) -> local CC -> send lists (owner
  keys, sort, dedup, segments) -> rep fill, sort, dedup, ghost extraction
  -> merge loop (init/hook/<=4-hop update) with the reference's exchange
  rule, a delta pack into pinned memory at every exchange -> labels.

At convergence the owned labels must equal the true global labels (the
minimum vertex id of each global component), which the caller supplies from
a full-graph solve; every pack can be checked against the oracle's pack.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native as nat
from .boundary import PinnedBuffer, delta_pack_device, delta_prepare_device, ghost_extract_device
from .cc import DeviceBuffer, DeviceGraph


def block_bounds(n: int, P: int) -> np.ndarray:
    """Vertex-block partition bounds (first n % P ranks one vertex more;
    src/graph_dist.c:22-51)."""
    base, rem = divmod(n, P)
    sizes = [base + 1 if r < rem else base for r in range(P)]
    return np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)


class RankShare:
    def __init__(self, full: DeviceGraph, P: int = 4, rank: int = 0):
        self.full, self.P, self.rank = full, P, rank
        self.bounds = block_bounds(full.n, P)
        self.vs, self.ve = int(self.bounds[rank]), int(self.bounds[rank + 1])
        self.nl = self.ve - self.vs
        self.times: dict[str, float] = {}
        self.stats: dict[str, int] = {}

    def _t(self, name, t0):
        nat.check(nat.gpu_lib().cc_hip_device_sync(), "sync")
        self.times[name] = self.times.get(name, 0.0) + (time.perf_counter() - t0)

    def preprocess(self) -> None:
        lib = nat.gpu_lib()
        nl, vs, ve, n = self.nl, self.vs, self.ve, self.full.n
        # the owned rows are a window of the full CSR (offsets stay global,
        # so the column base is the full graph's)
        d_row = self.full.row_ptr.ptr + 8 * vs
        d_col = self.full.col_idx.ptr
        t0 = time.perf_counter()
        self.lrow = DeviceBuffer(8 * (nl + 1))
        boff = DeviceBuffer(8 * (nl + 1))
        mll, nb = ctypes.c_uint64(), ctypes.c_uint64()
        nat.check(lib.cc_hip_split_count_device(d_row, d_col, nl, vs, ve, n, ctypes.byref(mll),
                                                ctypes.byref(nb), None), "split count")
        self.m_ll, self.nb = int(mll.value), int(nb.value)
        self.lcol = DeviceBuffer(max(4, 4 * self.m_ll))
        rec = DeviceBuffer(max(8, 8 * self.nb))
        nat.check(lib.cc_hip_split_emit_device(d_row, d_col, nl, vs, ve, n, self.lrow.ptr, boff.ptr,
                                               self.lcol.ptr, rec.ptr, None), "split emit")
        boff.close()
        self._t("split", t0)

        t0 = time.perf_counter()
        self.comp_of = DeviceBuffer(4 * max(1, nl))
        nat.check(lib.cc_hip_components_device(self.lrow.ptr, self.lcol.ptr, nl, self.m_ll,
                                               self.comp_of.ptr, None), "local CC")
        self._t("local_cc", t0)
        self.lrow.close()
        self.lcol.close()

        # send lists: the owned vertices each other rank has as ghosts (one
        # bitmask pass over the raw records, then one ordered compaction per
        # destination)
        t0 = time.perf_counter()
        d_bounds = DeviceBuffer.from_array(self.bounds)
        self.send_to_flat = DeviceBuffer(4 * max(1, min(self.nb, nl * self.P)))
        starts = np.zeros(self.P + 1, dtype=np.int32)
        nat.check(lib.cc_hip_send_lists_device(rec.ptr, self.nb, nl, d_bounds.ptr, self.P, vs,
                                               nat.ptr(starts, ctypes.c_int),
                                               self.send_to_flat.ptr, None), "send lists")
        nsend = int(starts[self.P])
        self.total_send = nsend
        # destination segments in dests order (ranks with something to send)
        self.dests = [q for q in range(self.P) if starts[q + 1] > starts[q]]
        self.seg_start = np.array([starts[q] for q in self.dests] + [nsend], dtype=np.int32)
        self._t("send_lists", t0)
        keys = DeviceBuffer(max(8, 8 * self.nb))

        # boundary records: rep fill, sort and dedup in one call (the most
        # frequent rep's records through a bitmap over the remotes, the rest
        # sorted by rank among the local roots, the two merged), then ghost
        # extraction
        t0 = time.perf_counter()
        kept, in_alt = ctypes.c_uint64(), ctypes.c_int()
        nat.check(lib.cc_hip_boundary_build_device(rec.ptr, keys.ptr, self.nb, self.comp_of.ptr, nl,
                                                   ctypes.byref(kept), ctypes.byref(in_alt), None),
                  "boundary build")
        self.nbd = int(kept.value)
        srt, oth = (keys, rec) if in_alt.value else (rec, keys)
        self._t("boundary_build", t0)
        t0 = time.perf_counter()
        self.ghosts = DeviceBuffer(4 * max(1, self.nbd))
        self.ng = ghost_extract_device(srt, self.nbd, oth, self.ghosts)
        self.bnd = oth
        self._t("ghost_extract", t0)
        for b in (srt, d_bounds):
            b.close()
        # comp_of is fixed from here on: the packs read the send entries'
        # reps from a list gathered once
        t0 = time.perf_counter()
        self.rep_of_send = DeviceBuffer(4 * max(1, nsend))
        delta_prepare_device(self.send_to_flat, nsend, self.comp_of, vs, ve, self.rep_of_send)
        self._t("delta_prepare", t0)
        # the rep list: the merge round's per-rep passes walk it instead of
        # scanning comp_of (src/cc_mpi.c:328-362)
        t0 = time.perf_counter()
        self.reps = DeviceBuffer(4 * max(1, nl))
        rep_count = ctypes.c_uint32(0)
        nat.check(lib.cc_hip_local_reps_device(self.comp_of.ptr, nl, self.reps.ptr,
                                               ctypes.byref(rep_count), None), "local reps")
        self.rep_count = int(rep_count.value)
        # the merge rounds' records, split once: the most frequent rep's as
        # a bitmap over the ghosts, the rest as a side list
        self.dom_bits = DeviceBuffer(4 * max(1, (self.ng + 31) // 32))
        self.side = DeviceBuffer(8 * max(1, self.nbd))
        n_side, dom = ctypes.c_uint64(), ctypes.c_uint32()
        nat.check(lib.cc_hip_merge_prepare_device(self.bnd.ptr, self.nbd, self.ng, self.reps.ptr,
                                                  self.rep_count, self.dom_bits.ptr,
                                                  self.side.ptr, ctypes.byref(n_side),
                                                  ctypes.byref(dom), None),
                  "merge prepare")
        self.n_side, self.dominant = int(n_side.value), int(dom.value)
        self._t("merge_prepare", t0)
        self.stats.update(n_local=nl, m_local_local=self.m_ll, boundary_raw=self.nb,
                          boundary_dedup=self.nbd, ghosts=self.ng, total_send_to=nsend,
                          outdegree=len(self.dests), side_records=self.n_side)

    def owner_local_labels(self) -> np.ndarray:
        """For every vertex, the label its owning rank holds right after its
        own local CC (v_start_q + comp_of_q[v - v_start_q]): what that rank
        sends at its first exchange.  Computed with the same split and
        local-CC kernels for every rank's block."""
        lib = nat.gpu_lib()
        out = np.empty(self.full.n, dtype=np.uint32)
        d_col = self.full.col_idx.ptr
        for q in range(self.P):
            a, b = int(self.bounds[q]), int(self.bounds[q + 1])
            nl = b - a
            if nl == 0:
                continue
            lrow, boff = DeviceBuffer(8 * (nl + 1)), DeviceBuffer(8 * (nl + 1))
            mll, nb = ctypes.c_uint64(), ctypes.c_uint64()
            d_row = self.full.row_ptr.ptr + 8 * a
            nat.check(lib.cc_hip_split_count_device(d_row, d_col, nl, a, b, self.full.n,
                                                    ctypes.byref(mll), ctypes.byref(nb), None),
                      "split count")
            lcol, rec = DeviceBuffer(max(4, 4 * mll.value)), DeviceBuffer(max(8, 8 * nb.value))
            nat.check(lib.cc_hip_split_emit_device(d_row, d_col, nl, a, b, self.full.n, lrow.ptr,
                                                   boff.ptr, lcol.ptr, rec.ptr, None), "split emit")
            comp = DeviceBuffer(4 * nl)
            nat.check(lib.cc_hip_components_device(lrow.ptr, lcol.ptr, nl, mll.value, comp.ptr,
                                                   None), "local CC")
            part = np.empty(nl, dtype=np.uint32)
            comp.download(part)
            out[a:b] = part + np.uint32(a)
            for buf in (lrow, boff, lcol, rec, comp):
                buf.close()
        return out

    def prepare_schedule(self, truth_labels: DeviceBuffer, seed: int) -> dict:
        """The injected deltas, built once per seed and kept on the device:
        ghost g receives min(g's id, its owner's local-CC label) at exchange
        mid[g] (when that comes before due[g]) and its true global label at
        exchange due[g] = geometric(1/2) - 1, mid = due // 2; both values
        are ids of vertices in g's component, as every real delta is.  The
        pairs (ghost index << 32) | label are sorted by exchange, so
        exchange k applies the slice start[k]..start[k+1]."""
        cache = getattr(self, "_sched", None)
        if cache is not None and cache["seed"] == seed:
            return cache
        if cache is not None:
            for b in ("pairs", "iota"):
                cache[b].close()
        lib = nat.gpu_lib()
        ng = self.ng
        ghost_ids = np.empty(ng, dtype=np.uint32)
        self.ghosts.download(ghost_ids)
        truth = np.empty(ng, dtype=np.uint32)
        tmp = DeviceBuffer(4 * max(1, ng))
        nat.check(lib.cc_hip_gather_u32_device(truth_labels.ptr, self.ghosts.ptr, ng, tmp.ptr, None),
                  "gather")
        tmp.download(truth)
        tmp.close()
        rng = np.random.default_rng(seed)
        due = rng.geometric(0.5, ng).astype(np.int64) - 1
        mid = due // 2
        if not hasattr(self, "_owner_local"):
            self._owner_local = self.owner_local_labels()
        half = np.flatnonzero(mid < due)
        gidx = np.concatenate([np.arange(ng, dtype=np.uint64), half.astype(np.uint64)])
        val = np.concatenate([truth, np.minimum(ghost_ids[half],
                                                self._owner_local[ghost_ids[half]])])
        # exchange indices are small (geometric(1/2) draws), so a stable
        # sort of them as bytes is a linear radix sort
        assert ng == 0 or int(due.max()) < 256
        when = np.concatenate([due, mid[half]]).astype(np.uint8)
        order = np.argsort(when, kind="stable")
        pairs = (gidx[order] << np.uint64(32)) | val[order].astype(np.uint64)
        last = int(due.max()) if ng else -1
        start = np.concatenate([[0], np.cumsum(np.bincount(when, minlength=last + 1))]).astype(np.int64)
        d_pairs = DeviceBuffer.from_array(pairs) if pairs.size else DeviceBuffer(8)
        iota = DeviceBuffer(4 * max(1, ng))
        nat.check(lib.cc_hip_iota_u32_device(iota.ptr, ng, 0, None), "iota")
        self._sched = dict(seed=seed, pairs=d_pairs, start=start, last=last, iota=iota,
                           zero=np.zeros(1, dtype=np.int32))
        return self._sched

    def merge(self, truth_labels: DeviceBuffer, exchange_interval: int = 1, seed: int = 1,
              check_pack=None, max_iters: int = 100000, use_reps: bool = True,
              prepared: bool = True) -> np.ndarray:
        """Run the merge loop with injected ghost deltas; returns the owned
        labels (comp_label[comp_of[li]]).  check_pack(seg_start, comp_label,
        comp_of, prev_before, pairs, counts, displs) is called per exchange
        if given.  use_reps: the packs read the reps gathered once after
        local CC (else each pack gathers them from comp_of).  prepared: the
        rounds run over the split records and keep each rep's minimum ghost
        label until the ghost labels change (else every round reads the
        whole record list)."""
        lib = nat.gpu_lib()
        nl, vs, ve, ng = self.nl, self.vs, self.ve, self.ng
        interval = max(1, exchange_interval)
        comp_label = DeviceBuffer(4 * max(1, nl))
        comp_min = DeviceBuffer(4 * max(1, nl))
        nat.check(lib.cc_hip_iota_u32_device(comp_label.ptr, nl, vs, None), "iota")
        reps, rep_count = self.reps, ctypes.c_uint32(self.rep_count)
        ghost_labels = DeviceBuffer(4 * max(1, ng))
        ghost_min = DeviceBuffer(4 * max(1, self.rep_count))
        ghosts_changed = 2  # first round: seed the per-rep ghost minimum
        nat.check(lib.cc_hip_memcpy_d2d(ghost_labels.ptr, self.ghosts.ptr, 4 * ng, None), "copy")
        sched = self.prepare_schedule(truth_labels, seed)
        last_exchange = sched["last"]
        prev_sent = DeviceBuffer(4 * max(1, self.total_send))
        ones = np.full(max(1, self.total_send), 0xFFFFFFFF, dtype=np.uint32)
        prev_sent.upload(ones[: max(1, self.total_send)])
        sendbuf = PinnedBuffer(max(1, self.total_send))
        comp_of_h = None
        if check_pack is not None:
            comp_of_h = np.empty(nl, dtype=np.uint32)
            self.comp_of.download(comp_of_h)
        pending_since = 1 if (ng and self.nbd and nl) else 0
        steps, k, iters, packs, emitted_total = 0, 0, 0, 0, 0
        t_merge = t_pack = 0.0
        t_loop0 = time.perf_counter()  # the loop alone, after the buffers above are set up
        changed = ctypes.c_int(0)
        while True:
            t0 = time.perf_counter()
            if prepared:
                nat.check(lib.cc_hip_merge_step_prepared_device(
                    self.dom_bits.ptr, self.dominant, self.side.ptr, self.n_side, self.ghosts.ptr,
                    ghost_labels.ptr, ng, self.comp_of.ptr, comp_label.ptr, comp_min.ptr,
                    ghost_min.ptr, nl, reps.ptr, rep_count.value, vs, ve, ghosts_changed,
                    ctypes.byref(changed), None), "merge step")
                ghosts_changed = 0
            else:
                nat.check(lib.cc_hip_merge_step_reps_device(
                    self.bnd.ptr, self.nbd, self.ghosts.ptr, ghost_labels.ptr, ng,
                    self.comp_of.ptr, comp_label.ptr, comp_min.ptr, nl, reps.ptr,
                    rep_count.value, vs, ve, ctypes.byref(changed), None), "merge step")
            t_merge += time.perf_counter() - t0
            iters += 1
            if changed.value:
                pending_since = 1
            others_pending = k <= last_exchange
            if not (pending_since or others_pending) or iters >= max_iters:
                break
            steps += 1
            # src/cc_mpi.c:669 exchanges when global_changed_any, the OR of
            # every rank's changed flag, is 0; the other ranks keep changing
            # while deltas from them remain scheduled
            global_changed_any = bool(changed.value) or others_pending
            if steps >= interval or not global_changed_any:
                if self.dests:
                    before = None
                    if check_pack is not None:
                        before = np.empty(self.total_send, dtype=np.uint32)
                        prev_sent.download(before)
                    t0 = time.perf_counter()
                    pairs, counts, displs = delta_pack_device(
                        self.send_to_flat, self.total_send, self.seg_start, comp_label,
                        self.comp_of, vs, ve, prev_sent, sendbuf,
                        d_rep_of_send=self.rep_of_send if use_reps else None)
                    emitted_total += int(pairs.size)
                    packs += 1
                    t_pack += time.perf_counter() - t0
                    if check_pack is not None:
                        cl = np.empty(nl, dtype=np.uint32)
                        comp_label.download(cl)
                        stf = np.empty(self.total_send, dtype=np.uint32)
                        self.send_to_flat.download(stf)
                        check_pack(stf, self.seg_start, cl, comp_of_h, vs, ve, before,
                                   pairs.copy(), counts, displs)
                # deltas from the other ranks for this exchange, applied by
                # the receive-side unpack kernel as received pairs would be
                lo = hi = 0
                if k <= last_exchange:
                    lo, hi = int(sched["start"][k]), int(sched["start"][k + 1])
                if hi > lo:
                    nat.check(lib.cc_hip_delta_unpack_device(
                        sched["pairs"].ptr + 8 * lo, hi - lo, nat.ptr(sched["zero"], ctypes.c_int),
                        nat.ptr(sched["zero"], ctypes.c_int), 1, sched["iota"].ptr,
                        ghost_labels.ptr, None), "delta unpack")
                    ghosts_changed = 1
                k += 1
                pending_since = 0
                steps = 0
        t_loop = time.perf_counter() - t_loop0
        labels = DeviceBuffer(4 * max(1, nl))
        roots = ctypes.c_uint64()
        nat.check(lib.cc_hip_rank_labels_device(comp_label.ptr, self.comp_of.ptr, nl, vs,
                                                labels.ptr, ctypes.byref(roots), None), "labels")
        out = np.empty(nl, dtype=np.uint32)
        labels.download(out)
        for b in (labels, comp_label, comp_min, ghost_labels, ghost_min, prev_sent):
            b.close()
        sendbuf.close()
        self.times["merge_steps"] = t_merge
        self.times["merge_loop"] = t_loop
        self.times["delta_packs"] = t_pack
        self.stats.update(merge_iterations=iters, exchanges=k, packs=packs,
                          pairs_emitted=emitted_total, local_roots=int(roots.value))
        return out

    def close(self) -> None:
        sched = getattr(self, "_sched", None)
        if sched is not None:
            sched["pairs"].close()
            sched["iota"].close()
            self._sched = None
        for name in ("comp_of", "send_to_flat", "ghosts", "bnd", "rep_of_send", "dom_bits",
                     "side", "reps"):
            b = getattr(self, name, None)
            if b is not None:
                b.close()
