"""Seeded workload generators: determinism and the shape each config asks for."""
import numpy as np

from helpers import packed
from sparkalog_amd import workload


def test_random_dag_shape(expected):
    s, d = workload.random_dag()
    assert s.size == 4000
    assert np.all(s < d) and np.all(d < 1000)           # i -> j with i < j
    k = packed((s, d))
    assert np.all(k[1:] > k[:-1])                        # sorted, duplicates removed
    assert workload.pair_digest(s, d) == expected["random_dag_1000_4000"]["edge_digest"]
    s2, d2 = workload.random_dag()
    assert np.array_equal(s, s2) and np.array_equal(d, d2)


def test_component_graph_digests(expected):
    for nodes in (1 << 14, 1 << 17):
        s, d, comps = workload.component_graph(nodes=nodes)
        e = expected[f"component_graph_{nodes}"]
        assert s.size == e["edges"] and comps == e["components"]
        assert workload.pair_digest(s, d) == e["edge_digest"]
        k = packed((s, d))
        assert np.all(k[1:] > k[:-1])
        assert s.max() < nodes and d.max() < nodes
        # Poisson(2) out-degree: about 2 edges per node after duplicate removal
        assert 1.8 < s.size / nodes < 2.05


def test_batch_graphs_shape():
    s, d, offs, counts = workload.batch_graphs(instances=64, n_max=2048)
    assert offs.size == 65 and offs[-1] == s.size
    assert counts.min() >= 16 and counts.max() <= 2048
    for i in range(64):
        a, b = s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]]
        assert a.size == 0 or (a.max() < counts[i] and b.max() < counts[i])


def test_operator_inputs():
    a, b = workload.distinct_input(5, 100000)
    uniq = np.unique(packed((a, b)))
    assert uniq.size == 70000
    assert a.max() < (1 << 24)
    (lf, ls), (rf, rs) = workload.overlapping_sets(6, 50000, 0.5)
    lk, rk = packed((lf, ls)), packed((rf, rs))
    assert np.all(lk[1:] > lk[:-1]) and np.all(rk[1:] > rk[:-1])
    assert np.intersect1d(lk, rk).size == 25000
    z, _ = workload.pairs(7, 200000, workload.KEY_ZIPF)
    counts = np.bincount(z[z < 1000], minlength=1000)
    assert counts[0] > counts[10] > counts[500]          # Zipf: rank 1 is the most frequent


def test_ordered_pair_digest():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 1 << 32, 1000, dtype=np.uint64).astype(np.uint32)
    b = rng.integers(0, 1 << 32, 1000, dtype=np.uint64).astype(np.uint32)
    d = workload.ordered_pair_digest(a, b)
    # one row: the same term as the order-independent digest's
    assert workload.ordered_pair_digest(a[:1], b[:1]) == workload.pair_digest(a[:1], b[:1])
    # two rows swapped: a different digest; chunking does not matter
    a2, b2 = a.copy(), b.copy()
    a2[[10, 500]], b2[[10, 500]] = a2[[500, 10]], b2[[500, 10]]
    assert workload.ordered_pair_digest(a2, b2) != d
    assert workload.ordered_pair_digest(a, b, chunk=7) == d


def test_overlapping_sets_with_zipf_keys():
    (lf, ls), (rf, rs) = workload.overlapping_sets(8, 40000, 0.5, workload.KEY_ZIPF)
    lk, rk = packed((lf, ls)), packed((rf, rs))
    assert np.all(lk[1:] > lk[:-1]) and np.all(rk[1:] > rk[:-1])
    assert np.intersect1d(lk, rk).size == 20000
    counts = np.bincount(lf[lf < 100], minlength=100)
    assert counts[0] > counts[50]


def test_plan_waves_respects_budget_and_id_space():
    from sparkalog_amd import batch

    nodes = np.full(100, 1 << 13, np.int64)   # 13 local bits: at most 2^19 members
    edges = np.full(100, 1 << 14, np.int64)
    waves = batch.plan_waves(nodes, edges, budget_pairs=1e30)
    assert waves == [list(range(100))]
    waves = batch.plan_waves(nodes, edges, budget_pairs=1e30, max_members=7)
    assert [len(w) for w in waves] == [7] * 14 + [2]
    assert sum(waves, []) == list(range(100))
    est = batch.estimate_closure(1 << 13, 1 << 14)
    waves = batch.plan_waves(nodes, edges, budget_pairs=3.5 * est)
    assert all(len(w) == 3 for w in waves[:-1]) and sum(waves, []) == list(range(100))
    # 2^18 local ids leave 2^14 members
    big = np.full(20000, 1 << 18, np.int64)
    waves = batch.plan_waves(big, np.ones(20000, np.int64), budget_pairs=1e30)
    assert max(len(w) for w in waves) == 1 << 14


def test_plan_waves_in_diameter_order_covers_every_instance_once():
    from sparkalog_amd import batch

    rng = np.random.default_rng(5)
    nodes = np.exp(rng.uniform(np.log(16), np.log(16384), 500)).astype(np.int64)
    edges = (nodes * rng.uniform(1.2, 3.0, 500)).astype(np.int64)
    order = batch.diameter_order(nodes, edges)
    assert sorted(order) == list(range(500))
    est = [batch.estimated_iterations(int(n), int(e)) for n, e in zip(nodes, edges)]
    assert all(est[a] >= est[b] for a, b in zip(order, order[1:]))
    budget = 5e8
    waves = batch.plan_waves(nodes, edges, budget_pairs=budget, order=order)
    assert sorted(sum(waves, [])) == list(range(500))
    for w in waves:
        total = sum(batch.estimate_closure(int(nodes[i]), int(edges[i])) for i in w)
        assert len(w) == 1 or total <= budget
