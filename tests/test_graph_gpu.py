"""Graph-replayed iterations (store option "graph") on the GPU: the same
results as the oracle, BFS and the regular executor, for transitive closure
(exact arrays), with the iteration captured once per buffer layout and then
replayed, through candidate and FULL buffer overflows (which grow the
buffers and run the iteration again), wide ids (hashed index), many short
iterations (a long chain), per-instance tracking (batched waves) and served
lanes."""
import numpy as np
import pytest

import oracle
from helpers import chain
from sparkalog_amd import batch, plans, serving, workload
from sparkalog_amd.recursion import RecursiveExecutor, RelationStore

pytestmark = pytest.mark.gpu
R = plans.RelationId


def _tc(edges, stream, graph=True):
    store = RelationStore(stream)
    store.set_graph(graph)
    store.insert_static(R(0), edges)
    store.insert_recursive(R(1), edges, edges)
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 100_000)
    return store, summ


def test_graph_dag_matches_oracle(expected, stream):
    e = expected["random_dag_1000_4000"]
    edges = workload.random_dag()
    store, summ = _tc(edges, stream)
    ref = oracle.Store()
    ref.insert_static(0, edges)
    ref.insert_recursive(1, edges, edges)
    rc, ref_summary = ref.run([1], [oracle.rule_from_plan(plans.transitive_closure_scc(R(1), R(0)).rules[0])],
                              10_000, 4)
    assert rc == 0
    full, exp = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], exp[0]) and np.array_equal(full[1], exp[1])
    assert summ.iterations == ref_summary.iterations == e["max_dist"]
    assert summ.total_new_rows == ref_summary.total_new_rows
    captures, replays = store.graph_stats()
    assert captures >= 1 and replays >= 1 and captures + replays >= summ.iterations


def test_graph_component_graph_with_overflows(expected, stream):
    # FULL grows from 3.2e4 to 1.47e7 rows: the FULL and candidate buffers
    # overflow several times, each time grown and the iteration run again
    e = expected["component_graph_16384"]
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store, summ = _tc((s, d), stream)
    bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
    full = store.full(R(1))
    assert np.array_equal(full[0], bfs["closure"][0]) and np.array_equal(full[1], bfs["closure"][1])
    assert summ.iterations == e["max_dist"]
    assert summ.total_new_rows == e["path_rows"] - e["edges"]
    # the regular executor on the same store type agrees on the summary
    _, plain = _tc((s, d), stream, graph=False)
    assert (plain.iterations, plain.total_new_rows) == (summ.iterations, summ.total_new_rows)


def test_graph_hands_over_to_the_regular_executor(expected, stream):
    # FULL (3.2e4 edges at first, 1.47e7 rows at the end) outgrows
    # graph_max_rows (2^18) after a few graph iterations: the rest of the
    # fixpoint runs in the regular executor, with the same result
    e = expected["component_graph_16384"]
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store = RelationStore(stream)
    store.set_graph(True)
    store.set_option("graph_max_rows", 1 << 18)
    store.insert_static(R(0), (s, d))
    store.insert_recursive(R(1), (s, d), (s, d))
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 100_000)
    digest, sorted_unique, rows = store.digest(R(1))
    assert (rows, digest, sorted_unique) == (e["path_rows"], e["path_digest"], True)
    assert summ.iterations == e["max_dist"]
    captures, replays = store.graph_stats()
    assert captures >= 1 and captures + replays < summ.iterations


@pytest.mark.parametrize("n", [2, 3, 1000])
def test_graph_chain_closed_form(n, stream):
    store, summ = _tc(chain(n), stream)
    assert store.rows(R(1)) == n * (n - 1) // 2
    assert summ.iterations == n - 1
    digest, sorted_unique, rows = store.digest(R(1))
    assert sorted_unique and rows == n * (n - 1) // 2


def test_graph_wide_ids(stream):
    nodes = 1 << 11
    s, d, _ = workload.component_graph(nodes=nodes)
    labels = workload.wide_labels(nodes)
    edges = workload.relabel(s, d, labels)
    store, summ = _tc(edges, stream)
    bfs = oracle.tc_bfs(s, d, nodes, closure=True)
    a, b = bfs["closure"]
    exp = np.sort((labels[a].astype(np.uint64) << np.uint64(32)) | labels[b].astype(np.uint64))
    full = store.full(R(1))
    assert np.array_equal((full[0].astype(np.uint64) << np.uint64(32)) | full[1], exp)
    assert summ.iterations == bfs["max_dist"]


def test_graph_falls_back_for_other_plans(golden, stream):
    # mutual recursion has two rules: the regular executor runs it
    from helpers import cols, pairs

    for case in golden["recursion"]:
        store = RelationStore(stream)
        store.set_graph(True)
        for rid, rel in case["static"].items():
            store.insert_static(R(int(rid)), cols(rel))
        for rid, st in case["recursive"].items():
            store.insert_recursive(R(int(rid)), cols(st["full"]), cols(st["delta"]))
        rules = tuple(plans.RecursiveRulePlan(target=R(r["target"]), delta_input=R(r["delta_input"]),
                                              right_input=R(r["right_input"]), join=plans.path_join_plan())
                      for r in case["rules"])
        summary = RecursiveExecutor.compile(plans.RecursiveSccPlan(tuple(R(r) for r in case["relations"]), rules),
                                            store).run(store, 16)
        if "iterations" in case:
            assert summary.iterations == case["iterations"], case["name"]
        for rid, rel in case["full"].items():
            assert pairs(store.full(R(int(rid)))) == [tuple(p) for p in rel], case["name"]


def test_graph_batched_matches_bfs(stream):
    s, d, offs, counts = workload.batch_graphs(seed=0x5B, instances=160, n_max=2048)
    store = RelationStore(stream)
    store.set_graph(True)
    total_est = sum(batch.estimate_closure(int(n), int(e)) for n, e in zip(counts, np.diff(offs.astype(np.int64))))
    res = batch.run_batched(s, d, offs, counts, store, budget_pairs=total_est / 3)
    for i in range(counts.size):
        r = oracle.tc_bfs(s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]], int(counts[i]), threads=8)
        assert (int(res.rows[i]), int(res.digests[i]), int(res.iterations[i])) == \
            (r["pairs"], r["digest"], r["max_dist"]), i


def test_graph_served_lanes_equal_alone(stream):
    import torch
    from sparkalog_amd.execution import HipStream

    reqs = serving.make_requests(30, rate_per_s=600.0, seed=0x5F, min_log2_edges=8, max_log2_edges=14)
    stores = [RelationStore(HipStream(torch.cuda.Stream())) for _ in range(3)]
    for st in stores:
        st.set_graph(True)
    mixed, _ = serving.serve_lanes(reqs, stores, bounds=[2e4, 2e5], budget_pairs=1e9, max_slots=64)
    plain = RelationStore(stream)
    plain.set_graph(True)
    alone, _ = serving.serve_plain(reqs, plain)
    for m, a, req in zip(mixed, alone, reqs):
        bfs = oracle.tc_bfs(req.src, req.dst, req.nodes)
        assert (m.rows, m.digest) == (a.rows, a.digest) == (bfs["pairs"], bfs["digest"]), req.rid
