"""Batched, continuously admitted CC serving on one GPU.

Requests name a graph; graphs live in a resident store in HBM (CSR uploaded
once, admitted against an HBM byte budget, least-recently-used graphs that
no request holds are evicted).  A server with `lanes` batch engines
(`cc_hip_batch_*`: one stream, device tables and one captured HIP graph of
the whole solve each) forms a batch from every request that has arrived as
soon as a lane is free, within the lane's vertex/item capacity, launches it
as one graph replay, and queues the label copies to pinned host memory on
the same stream; requests that arrive meanwhile go into the next lane's
batch, so their work overlaps on the device.  A batch of one request (more
generally, of at most `direct_max`) skips the replay and runs the
single-graph solve on the lane's stream, so light load pays no replay
overhead.  Latency = completion of the label copy - arrival.

The one-graph-per-launch baseline serves the same arrival schedule FIFO:
one `cc_hip_components_device` solve and one label copy per request.

Labels never depend on the batch a request lands in: every graph is solved
in its own label array, and the result is the minimum vertex id of each
component of that graph.
"""
from __future__ import annotations

import collections
import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _native as nat
from .cc import DeviceBuffer, DeviceGraph


def _lib():
    lib = nat.gpu_lib()
    if not hasattr(lib, "_serving_sigs"):
        P = ctypes.c_void_p
        lib.cc_hip_batch_create.restype = ctypes.c_int
        lib.cc_hip_batch_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.POINTER(P)]
        lib.cc_hip_batch_set.restype = ctypes.c_int
        lib.cc_hip_batch_set.argtypes = [ctypes.c_char_p, ctypes.c_int]
        lib.cc_hip_batch_destroy.restype = ctypes.c_int
        lib.cc_hip_batch_destroy.argtypes = [P]
        lib.cc_hip_batch_launch.restype = ctypes.c_int
        lib.cc_hip_batch_launch.argtypes = [P, ctypes.c_uint32, ctypes.POINTER(P), ctypes.POINTER(P),
                                            nat.P_u32, ctypes.POINTER(P), ctypes.c_uint32,
                                            nat.P_u32]
        lib.cc_hip_batch_query.restype = ctypes.c_int
        lib.cc_hip_batch_query.argtypes = [P]
        lib.cc_hip_batch_sync.restype = ctypes.c_int
        lib.cc_hip_batch_sync.argtypes = [P]
        lib.cc_hip_batch_stream.restype = P
        lib.cc_hip_batch_stream.argtypes = [P]
        lib.cc_hip_memcpy_d2h_async.restype = ctypes.c_int
        lib.cc_hip_memcpy_d2h_async.argtypes = [P, P, ctypes.c_size_t, P]
        lib._serving_sigs = True
    return lib


class PinnedU32:
    """Pinned host uint32 array."""

    def __init__(self, count: int):
        p = ctypes.c_void_p()
        nat.check(_lib().cc_hip_host_alloc(ctypes.byref(p), max(4, 4 * count)), "cc_hip_host_alloc")
        self.ptr = p.value
        buf = (ctypes.c_char * max(4, 4 * count)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=np.uint32, count=count)

    def close(self):
        if self.ptr:
            self.array = None
            _lib().cc_hip_host_free(self.ptr)
            self.ptr = None


class BatchEngine:
    """One lane: a cc_hip_batch engine plus device/pinned label arenas."""

    def __init__(self, max_vertices: int, max_items: int, max_graphs: int):
        lib = _lib()
        h = ctypes.c_void_p()
        nat.check(lib.cc_hip_batch_create(max_items, max_graphs, ctypes.byref(h)),
                  "cc_hip_batch_create")
        self.handle = h.value
        self.stream = lib.cc_hip_batch_stream(self.handle)
        self.max_vertices, self.max_items, self.max_graphs = max_vertices, max_items, max_graphs
        self.labels = DeviceBuffer(4 * max_vertices)
        self.out = PinnedU32(max_vertices)
        self.direct_launches = 0

    def launch(self, graphs: list[DeviceGraph], item_vertices: int,
               direct_max: int = 0) -> list[tuple[int, int]]:
        """Solve `graphs` on this lane's stream; returns each graph's (offset,
        n) in the label arenas.  Label copies to pinned memory are queued
        behind the solve.  A batch of at most `direct_max` graphs runs as
        that many single-graph solves (`cc_hip_components_device`) instead
        of one replay: a replay carries fixed work (table upload, every
        stage of the schedule) that a lone small request does not amortise."""
        lib = _lib()
        k = len(graphs)
        offs, off = [], 0
        for g in graphs:
            offs.append((off, g.n))
            off += g.n
        if off > self.max_vertices:
            raise ValueError("batch exceeds the lane's vertex capacity")
        if k <= direct_max:
            for g, (o, _) in zip(graphs, offs):
                nat.check(lib.cc_hip_components_device(g.row_ptr.ptr, g.col_idx.ptr, g.n, g.m,
                                                       self.labels.ptr + 4 * o, self.stream),
                          "cc_hip_components_device")
            self.direct_launches += 1
        else:
            P = ctypes.c_void_p
            rp = (P * k)(*[g.row_ptr.ptr for g in graphs])
            ci = (P * k)(*[g.col_idx.ptr for g in graphs])
            nn = (ctypes.c_uint32 * k)(*[g.n for g in graphs])
            lb = (P * k)(*[self.labels.ptr + 4 * o for o, _ in offs])
            used = ctypes.c_uint32(0)
            nat.check(lib.cc_hip_batch_launch(self.handle, k, rp, ci, nn, lb, item_vertices,
                                              ctypes.byref(used)), "cc_hip_batch_launch")
        if off:
            nat.check(lib.cc_hip_memcpy_d2h_async(self.out.ptr, self.labels.ptr, 4 * off,
                                                  self.stream), "label copy")
        return offs

    def done(self) -> bool:
        r = _lib().cc_hip_batch_query(self.handle)
        if r < 0:
            nat.check(1, "cc_hip_batch_query")
        return r == 1

    def sync(self) -> None:
        nat.check(_lib().cc_hip_batch_sync(self.handle), "cc_hip_batch_sync")

    def items_for(self, graphs, item_vertices) -> int:
        return sum(math.ceil(g.n / item_vertices) for g in graphs)

    def close(self) -> None:
        if self.handle:
            _lib().cc_hip_batch_destroy(self.handle)
            self.handle = None
            self.labels.close()
            self.out.close()


def resident_bytes(g: DeviceGraph) -> int:
    """HBM a resident graph holds: its row offsets, column and label buffers."""
    return g.row_ptr.nbytes + g.col_idx.nbytes + g.labels.nbytes


class GraphStore:
    """Resident graphs in HBM under a byte budget.

    A graph is admitted only if its cost fits: `estimate(key)` gives
    (resident, peak) bytes from the key's spec before anything is built,
    where peak covers the loader's transient buffers; least-recently-used
    graphs that no in-flight request holds are evicted to make room, and a
    key that still does not fit is refused without being built and stays
    refused until some graph is released.  Without an estimate, the graph is
    loaded first and priced by what it holds (the transient build is then
    not bounded)."""

    def __init__(self, budget_bytes: int, loader, estimate=None):
        self.budget = budget_bytes
        self.loader = loader  # key -> DeviceGraph (builds/uploads it)
        self.estimate = estimate  # key -> (resident, peak) bytes, or None
        self.graphs: "collections.OrderedDict[object, DeviceGraph]" = collections.OrderedDict()
        self.cost: dict = {}
        self.pins: collections.Counter = collections.Counter()
        self.refused: set = set()
        self.used = 0
        self.peak_used = 0
        self.loads = self.evictions = self.refusals = 0

    def _make_room(self, need: int) -> bool:
        """Evict unpinned graphs (LRU first) until `need` more bytes fit;
        evicts nothing if even evicting every unpinned graph would not do."""
        if need > self.budget:
            return False
        free_able = sum(self.cost[k] for k in self.graphs if self.pins[k] == 0)
        if self.used - free_able + need > self.budget:
            return False
        while self.used + need > self.budget:
            victim = next(k for k in self.graphs if self.pins[k] == 0)
            self._evict(victim)
        return True

    def acquire(self, key) -> DeviceGraph | None:
        """The resident graph for key (loading it if there is room), or None
        if it does not fit until other graphs are released."""
        if key in self.graphs:
            self.graphs.move_to_end(key)
            self.pins[key] += 1
            return self.graphs[key]
        if key in self.refused:
            return None
        if self.estimate is not None:
            resident, peak = self.estimate(key)
            if not self._make_room(max(resident, peak)):
                self.refused.add(key)
                self.refusals += 1
                return None
            g = self.loader(key)
            cost = resident_bytes(g)
            if cost > resident:
                g.close()
                raise RuntimeError(f"graph {key!r} holds {cost} bytes, above its estimate "
                                   f"{resident}")
        else:
            g = self.loader(key)
            cost = resident_bytes(g)
            if not self._make_room(cost):
                g.close()
                self.refused.add(key)
                self.refusals += 1
                return None
        self.loads += 1
        self.graphs[key] = g
        self.cost[key] = cost
        self.pins[key] += 1
        self.used += cost
        self.peak_used = max(self.peak_used, self.used)
        return g

    def release(self, key) -> None:
        self.pins[key] -= 1
        self.refused.clear()

    def _evict(self, key) -> None:
        g = self.graphs.pop(key)
        self.used -= self.cost.pop(key)
        self.evictions += 1
        g.close()

    def stats(self) -> dict:
        return {"budget_bytes": self.budget, "resident_bytes": self.used,
                "peak_resident_bytes": self.peak_used, "loads": self.loads,
                "evictions": self.evictions, "refusals": self.refusals,
                "resident_graphs": len(self.graphs)}

    def close(self) -> None:
        for k in list(self.graphs):
            self._evict(k)


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the start of the run
    key: object
    done: float = math.nan
    labels: np.ndarray | None = field(default=None, repr=False)


def autotune_item_vertices(engine: BatchEngine, graphs: list[DeviceGraph],
                           candidates=(1024, 4096, 16384, 65536), reps: int = 5) -> dict:
    """Time a representative batch under each work-item size; returns the
    timings and the fastest size."""
    fit, verts = [], 0
    for g in graphs:  # the representative batch: as much of the set as one lane holds
        if verts + g.n <= engine.max_vertices and len(fit) < engine.max_graphs:
            fit.append(g)
            verts += g.n
    graphs = fit
    timings = {}
    for iv in candidates:
        if engine.items_for(graphs, iv) > engine.max_items:
            continue
        engine.launch(graphs, iv)
        engine.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            engine.launch(graphs, iv)
            engine.sync()
        timings[iv] = (time.perf_counter() - t0) / reps
    best = min(timings, key=timings.get)
    return {"item_vertices": best, "seconds": {str(k): round(v, 6) for k, v in timings.items()}}


class Server:
    """Continuous batching over `lanes` batch engines."""

    def __init__(self, store: GraphStore, lanes: int = 2, max_vertices: int = 1 << 25,
                 max_items: int = 1 << 15, max_graphs: int = 1024, item_vertices: int = 4096,
                 keep_labels: bool = False, direct_max: int = 1):
        self.store = store
        self.engines = [BatchEngine(max_vertices, max_items, max_graphs) for _ in range(lanes)]
        self.item_vertices = item_vertices
        self.keep_labels = keep_labels
        self.direct_max = direct_max  # batches this small run as single-graph solves

    def close(self):
        for e in self.engines:
            e.close()

    def run(self, requests: list[Request], on_result=None) -> dict:
        """Serve `requests` on their arrival schedule.  on_result(req, labels)
        is called for every request as its batch retires, with a view of its
        labels in the lane's pinned buffer (valid only during the call)."""
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        queue: collections.deque = collections.deque()
        busy: dict[int, list] = {}
        batches = direct = waits = 0
        t0 = time.perf_counter()
        total = len(pending)
        finished = 0
        while finished < total:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                queue.append(pending.popleft())
            # retire finished lanes
            for li in list(busy):
                eng = self.engines[li]
                if eng.done():
                    tdone = time.perf_counter() - t0
                    for req, g, (off, n) in busy.pop(li):
                        req.done = tdone
                        if self.keep_labels:
                            req.labels = eng.out.array[off:off + n].copy()
                        if on_result is not None:
                            on_result(req, eng.out.array[off:off + n])
                        self.store.release(req.key)
                        finished += 1
            # fill free lanes with everything that has arrived
            for li, eng in enumerate(self.engines):
                if li in busy or not queue:
                    continue
                chosen, graphs, verts, items = [], [], 0, 0
                while queue and len(chosen) < eng.max_graphs:
                    req = queue[0]
                    g = self.store.acquire(req.key)
                    if g is None:  # waits for HBM until a lane retires
                        waits += 1
                        break
                    it = math.ceil(g.n / self.item_vertices)
                    if chosen and (verts + g.n > eng.max_vertices or items + it > eng.max_items):
                        self.store.release(req.key)
                        break
                    if g.n > eng.max_vertices or it > eng.max_items:
                        self.store.release(req.key)
                        raise ValueError(f"request {req.rid} exceeds a lane's capacity")
                    queue.popleft()
                    chosen.append(req)
                    graphs.append(g)
                    verts += g.n
                    items += it
                if chosen:
                    offs = eng.launch(graphs, self.item_vertices, self.direct_max)
                    busy[li] = list(zip(chosen, graphs, offs))
                    batches += 1
                    direct += len(graphs) <= self.direct_max
            if queue and not busy:
                raise MemoryError(f"request {queue[0].rid} does not fit the graph store budget")
            if not busy and not queue and pending:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 0:
                    time.sleep(min(wait, 0.002))
        wall = time.perf_counter() - t0
        return {"batches": batches, "direct_batches": direct, "admission_waits": waits,
                "wall_s": wall}


def serve_one_by_one(store: GraphStore, requests: list[Request], keep_labels: bool = False,
                     on_result=None) -> dict:
    """Baseline: FIFO, one solve launch sequence and one label copy per
    request.  on_result(req, labels) as in Server.run."""
    lib = nat.gpu_lib()
    pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
    out = None
    t0 = time.perf_counter()
    while pending:
        req = pending.popleft()
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        g = store.acquire(req.key)
        if g is None:
            raise MemoryError("graph store budget too small for one request")
        if out is None or out.array.size < g.n:
            if out is not None:
                out.close()
            out = PinnedU32(max(g.n, 1 << 22))
        g.solve_async()
        nat.check(lib.cc_hip_memcpy_d2h_async(out.ptr, g.labels.ptr, 4 * g.n, None), "copy")
        nat.check(lib.cc_hip_stream_sync(None), "sync")
        req.done = time.perf_counter() - t0
        if keep_labels:
            req.labels = out.array[:g.n].copy()
        if on_result is not None:
            on_result(req, out.array[:g.n])
        store.release(req.key)
    if out is not None:
        out.close()
    return {"wall_s": time.perf_counter() - t0}


def latency_summary(requests: list[Request], entries: dict) -> dict:
    lat = np.array([r.done - r.arrival for r in requests]) * 1e3
    wall = max(r.done for r in requests)
    m_total = sum(entries[r.key] for r in requests)
    return {"requests": len(requests), "p50_ms": round(float(np.percentile(lat, 50)), 3),
            "p99_ms": round(float(np.percentile(lat, 99)), 3),
            "mean_ms": round(float(lat.mean()), 3),
            "throughput_rps": round(len(requests) / wall, 1),
            "throughput_gteps": round(m_total / wall / 1e9, 3)}


def make_pool(per_type: int, seed: int, n_min: int = 1024, n_max: int = 4194304):
    """The request graphs: `per_type` graphs of each type (Erdos-Renyi with
    average degree 8, R-MAT edge factor 16, 2-D grid with 30% of edges
    deleted), vertex counts log-uniform in [n_min, n_max], built in HBM.
    Returns (keys, loader, estimate) for a GraphStore: estimate(key) prices
    a graph from its spec, before it is built."""
    from . import devgen
    rng = np.random.default_rng(seed)
    specs = {}
    for t in ("er", "rmat", "grid"):
        for i in range(per_type):
            n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
            specs[(t, i)] = (n, seed * 1000 + len(specs))

    def shape(key):
        t, _ = key
        n, _ = specs[key]
        if t == "er":
            return ("er", n) + devgen.er_params(n, 8.0)
        if t == "rmat":
            return ("rmat", max(10, int(round(math.log2(n))))) + devgen.rmat_params(
                max(10, int(round(math.log2(n)))), 16)
        side = max(2, int(round(math.sqrt(n))))
        return ("grid", side) + devgen.grid_params(side, side, 0.7)

    def loader(key):
        kind, arg, _, _ = shape(key)
        s = specs[key][1]
        if kind == "er":
            return devgen.erdos_renyi(arg, 8.0, seed=s)
        if kind == "rmat":
            return devgen.rmat(arg, 16, seed=s)
        return devgen.grid(arg, arg, 0.7, seed=s)

    def estimate(key):
        kind, _, n, params = shape(key)
        resident, peak = devgen.footprint(kind, n, params, permute=kind != "er")
        # the library's per-stream scratch and allocation rounding
        return resident + (1 << 20), peak + peak // 16 + (16 << 20)

    return list(specs), loader, estimate


def poisson_requests(count: int, qps: float, keys: list, seed: int) -> list[Request]:
    """count requests with exponential inter-arrival times at rate qps, each
    naming a key drawn uniformly from `keys`."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, count)
    arrivals = np.cumsum(gaps) - gaps[0]
    picks = rng.integers(0, len(keys), count)
    return [Request(i, float(arrivals[i]), keys[picks[i]]) for i in range(count)]
