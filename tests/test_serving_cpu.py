"""Host-side serving logic without a GPU: the graph store's HBM-budget
admission and eviction, and the Poisson request stream."""
import numpy as np
import pytest

from cc_mpi_amd import serving


class _Buf:
    def __init__(self, nbytes):
        self.nbytes = nbytes


class FakeGraph:
    def __init__(self, n, m):
        self.n, self.m, self.closed = n, m, False
        self.row_ptr, self.col_idx, self.labels = _Buf(8 * (n + 1)), _Buf(4 * m), _Buf(4 * n)

    def close(self):
        self.closed = True


def cost(n, m):
    return 8 * (n + 1) + 4 * m + 4 * n


def make_store(budget, sizes, with_estimate=True, peak_extra=0):
    made = []

    def loader(key):
        g = FakeGraph(*sizes[key])
        made.append(g)
        return g

    def estimate(key):
        c = cost(*sizes[key])
        return c, c + peak_extra
    return serving.GraphStore(budget, loader, estimate if with_estimate else None), made


@pytest.mark.parametrize("with_estimate", [True, False])
def test_admission_and_lru_eviction(with_estimate):
    sizes = {k: (1000, 10000) for k in "abcd"}
    one = cost(1000, 10000)
    store, made = make_store(int(2.5 * one), sizes, with_estimate)
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
    st = store.stats()
    assert st["loads"] == 3 and st["evictions"] == 1 and st["refusals"] == 1
    store.close()
    assert store.used == 0 and not store.graphs


def test_estimate_refuses_before_building():
    """With an estimate, a graph that does not fit is never built, and the
    refusal is remembered (no rebuild per poll) until a release."""
    sizes = {"a": (1000, 10000), "b": (1000, 10000)}
    one = cost(1000, 10000)
    store, made = make_store(int(1.5 * one), sizes)
    assert store.acquire("a") is not None
    assert len(made) == 1
    for _ in range(5):
        assert store.acquire("b") is None
    assert len(made) == 1  # b was never built
    assert store.refusals == 1  # remembered, not re-estimated
    store.release("a")
    assert store.acquire("b") is not None  # a evicted, b built
    assert len(made) == 2 and made[0].closed and store.used == one
    store.close()


def test_build_peak_counts_against_budget():
    """The loader's transient peak must fit too, not only what stays resident."""
    sizes = {"a": (1000, 10000), "b": (1000, 10000)}
    one = cost(1000, 10000)
    store, made = make_store(2 * one + 10, sizes, peak_extra=one)
    assert store.acquire("a") is not None
    assert store.acquire("b") is None  # one + (one + one) > budget while a is pinned
    assert len(made) == 1
    store.release("a")
    assert store.acquire("b") is not None  # evicting a makes room for the peak
    assert store.evictions == 1 and store.used == one
    store.close()


def test_no_eviction_when_it_cannot_help():
    sizes = {"a": (1000, 10000), "big": (100000, 1000000)}
    store, made = make_store(3 * cost(1000, 10000), sizes)
    assert store.acquire("a") is not None
    store.release("a")
    assert store.acquire("big") is None  # larger than the whole budget
    assert "a" in store.graphs and store.evictions == 0


def test_estimate_below_actual_is_an_error():
    sizes = {"a": (1000, 10000)}

    def estimate(key):
        return 10, 10
    store = serving.GraphStore(1 << 30, lambda k: FakeGraph(*sizes[k]), estimate)
    with pytest.raises(RuntimeError, match="above its estimate"):
        store.acquire("a")


def test_request_larger_than_budget_is_never_admitted():
    store, _ = make_store(100, {"x": (1000, 10000)})
    assert store.acquire("x") is None
    assert store.used == 0


def test_pool_estimates_bound_the_device_builder():
    """The pool's estimate covers the builder's buffers for every pool graph
    (resident: row offsets, the pair buffer kept as columns, labels)."""
    from cc_mpi_amd import devgen
    keys, _, estimate = serving.make_pool(per_type=4, seed=3)
    for k in keys:
        resident, peak = estimate(k)
        assert peak >= resident > 0
    n, params = devgen.grid_params(100, 100, 0.7)
    res, peak = devgen.footprint("grid", n, params)
    assert res == 8 * (n + 1) + 8 * 2 * 100 * 100 + 4 * n
    assert peak >= 8 * 20000 + 32 * 20000


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
