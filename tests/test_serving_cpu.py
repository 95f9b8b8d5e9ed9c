"""Host-side serving logic without a GPU: the graph store's HBM-budget
admission and eviction, and the Poisson request stream."""
import numpy as np
import pytest

from cc_mpi_amd import serving


class FakeGraph:
    def __init__(self, n, m):
        self.n, self.m, self.closed = n, m, False

    def close(self):
        self.closed = True


def make_store(budget, sizes):
    made = []

    def loader(key):
        g = FakeGraph(*sizes[key])
        made.append(g)
        return g
    return serving.GraphStore(budget, loader), made


def test_admission_and_lru_eviction():
    sizes = {k: (1000, 10000) for k in "abcd"}
    one = serving.footprint_bytes(1000, 10000)
    store, made = make_store(int(2.5 * one), sizes)
    assert store.acquire("a") is not None
    assert store.acquire("b") is not None
    assert store.acquire("c") is None  # both residents pinned: no room
    assert store.used == 2 * one
    store.release("a")
    assert store.acquire("c") is not None  # evicts a (unpinned, least recent)
    assert "a" not in store.graphs and store.used == 2 * one
    assert any(g.closed for g in made)
    store.release("b")
    store.release("c")
    assert store.acquire("b") is not None  # resident: no load
    store.release("b")
    store.close()
    assert store.used == 0 and not store.graphs


def test_request_larger_than_budget_is_never_admitted():
    store, _ = make_store(100, {"x": (1000, 10000)})
    assert store.acquire("x") is None
    assert store.used == 0


def test_poisson_stream():
    reqs = serving.poisson_requests(20000, qps=1000.0, keys=["a", "b", "c"], seed=1)
    t = np.array([r.arrival for r in reqs])
    assert t[0] == 0 and (np.diff(t) >= 0).all()
    assert abs(len(reqs) / t[-1] - 1000.0) < 30.0  # rate
    assert {r.key for r in reqs} == {"a", "b", "c"}
    again = serving.poisson_requests(20000, qps=1000.0, keys=["a", "b", "c"], seed=1)
    assert [r.arrival for r in again] == [r.arrival for r in reqs]


def test_latency_summary():
    reqs = [serving.Request(i, float(i), "k", done=float(i) + 0.001 * (i + 1)) for i in range(100)]
    s = serving.latency_summary(reqs, {"k": 1000})
    assert s["requests"] == 100
    assert s["p50_ms"] == pytest.approx(50.5, abs=0.6)
    assert s["p99_ms"] == pytest.approx(99.0, abs=1.1)
