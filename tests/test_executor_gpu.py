"""Device-resident semi-naive executor on the GPU against the oracle:
the reference's recursion tests, plain TC on the 1,000-node DAG (exact
arrays), and single-instance TC on component graphs (exact arrays at 2^14
nodes; row count, order-independent digest, strict sortedness and iteration
count against brute-force BFS fixtures at 2^17 and 2^20 nodes)."""
import numpy as np
import pytest

import oracle
from helpers import chain, cols, pairs
from sparkalog_amd import plans, workload
from sparkalog_amd.recursion import GenericExecutionError, RecursiveExecutor, RelationStore

pytestmark = pytest.mark.gpu

R = plans.RelationId


def rule(target, delta, right):
    return plans.RecursiveRulePlan(target=R(target), delta_input=R(delta), right_input=R(right),
                                   join=plans.path_join_plan())


def test_reference_recursion_vectors(golden, stream):
    for case in golden["recursion"]:
        store = RelationStore(stream)
        for rid, rel in case["static"].items():
            store.insert_static(R(int(rid)), cols(rel))
        for rid, st in case["recursive"].items():
            store.insert_recursive(R(int(rid)), cols(st["full"]), cols(st["delta"]))
        plan = plans.RecursiveSccPlan(tuple(R(r) for r in case["relations"]),
                                      tuple(rule(r["target"], r["delta_input"], r["right_input"]) for r in case["rules"]))
        summary = RecursiveExecutor.compile(plan, store).run(store, 16)
        if "iterations" in case:
            assert summary.iterations == case["iterations"], case["name"]
        if "total_new_rows" in case:
            assert summary.total_new_rows == case["total_new_rows"], case["name"]
        for rid, rel in case["full"].items():
            assert pairs(store.full(R(int(rid)))) == [tuple(p) for p in rel], case["name"]
            assert store.rows(R(int(rid)), 1) == 0


def test_clause_order_independence(stream):
    def run(reverse):
        store = RelationStore(stream)
        store.insert_static(R(0), cols([(2, 3), (3, 4)]))
        store.insert_static(R(1), cols([(2, 3), (3, 4), (4, 5)]))
        store.insert_recursive(R(2), cols([(1, 2)]), cols([(1, 2)]))
        rules = [rule(2, 2, 0), rule(2, 2, 1)]
        if reverse:
            rules.reverse()
        s = RecursiveExecutor.compile(plans.RecursiveSccPlan((R(2),), tuple(rules)), store).run(store, 16)
        return s.iterations, s.total_new_rows, pairs(store.full(R(2)))
    assert run(False) == run(True) == (4, 3, [(1, 2), (1, 3), (1, 4), (1, 5)])


def test_compile_errors(stream):
    store = RelationStore(stream)
    store.insert_static(R(0), cols([(1, 2)]))
    store.insert_recursive(R(1), cols([(1, 2)]), cols([(1, 2)]))
    first = rule(1, 1, 0)
    bad = plans.RecursiveRulePlan(R(1), R(1), R(0), plans.path_join_plan(),
                                  union=plans.SortedBinaryUnion((0, 1), (1, 0)))
    with pytest.raises(GenericExecutionError) as err:
        RecursiveExecutor.compile(plans.RecursiveSccPlan((R(1),), (first, bad)), store)
    assert err.value.kind == "IncompatibleTargetPlans"
    with pytest.raises(GenericExecutionError) as err:
        RecursiveExecutor.compile(plans.RecursiveSccPlan((), (first,)), store)
    assert err.value.kind == "EmptyScc"
    with pytest.raises(GenericExecutionError) as err:
        RecursiveExecutor.compile(plans.RecursiveSccPlan((R(0),), (rule(0, 0, 0),)), store)
    assert err.value.kind == "NotRecursive"
    with pytest.raises(GenericExecutionError) as err:
        store.insert_static(R(0), cols([(5, 6)]))
    assert err.value.kind == "Duplicate"


def test_iteration_limit(stream):
    store = RelationStore(stream)
    e = chain(10)
    store.insert_static(R(0), e)
    store.insert_recursive(R(1), e, e)
    with pytest.raises(GenericExecutionError) as err:
        RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 3)
    assert err.value.kind == "IterationLimit"


def _tc(edges, stream, max_iterations=10_000, sparse_probe_rows=None):
    store = RelationStore(stream)
    if sparse_probe_rows is not None:
        store.set_tuning(sparse_probe_rows=sparse_probe_rows)
    store.insert_static(R(0), edges)
    store.insert_recursive(R(1), edges, edges)
    summary = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, max_iterations)
    return store, summary


@pytest.mark.parametrize("n", [2, 3, 64, 1000])
def test_chain_closed_form(n, stream):
    store, s = _tc(chain(n), stream)
    assert s.relation_rows == [(R(1), n * (n - 1) // 2)]
    assert s.iterations == n - 1


def test_plain_tc_exact(expected, stream):
    e = expected["random_dag_1000_4000"]
    edges = workload.random_dag()
    store, s = _tc(edges, stream)
    ref = oracle.Store()
    ref.insert_static(0, edges)
    ref.insert_recursive(1, edges, edges)
    rc, rs = ref.run([1], [oracle.rule_from_plan(rule(1, 1, 0))], 10_000, 4)
    assert rc == 0
    full = store.full(R(1))
    expect = ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert s.iterations == rs.iterations == e["max_dist"]
    assert s.total_new_rows == rs.total_new_rows == e["path_rows"] - e["edges"]
    assert workload.pair_digest(*full) == e["path_digest"]


def test_component_graph_exact_small(expected, stream):
    e = expected["component_graph_16384"]
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store, summ = _tc((s, d), stream)
    full = store.full(R(1))
    bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
    assert np.array_equal(full[0], bfs["closure"][0]) and np.array_equal(full[1], bfs["closure"][1])
    assert summ.iterations == e["max_dist"]
    assert summ.total_new_rows == e["path_rows"] - e["edges"]


@pytest.mark.parametrize("nodes", [1 << 17, 1 << 20])
def test_component_graph_digest(nodes, expected, stream):
    e = expected[f"component_graph_{nodes}"]
    s, d, _ = workload.component_graph(nodes=nodes)
    store, summ = _tc((s, d), stream)
    digest, sorted_unique, rows = store.digest(R(1))
    assert rows == e["path_rows"]
    assert sorted_unique
    assert digest == e["path_digest"]
    assert summ.iterations == e["max_dist"]
    assert summ.total_new_rows == e["path_rows"] - e["edges"]


def _oracle_run(static, recursive, rules, relations, max_iterations=10_000):
    ref = oracle.Store()
    for rid, rel in static.items():
        ref.insert_static(rid, rel)
    for rid, (full, delta) in recursive.items():
        ref.insert_recursive(rid, full, delta)
    rc, summ = ref.run(relations, [oracle.rule_from_plan(r) for r in rules], max_iterations, 4)
    assert rc == 0
    return ref, summ


def test_long_candidate_runs(stream):
    # source 0 reaches 6,000 nodes in one step, each with 4 successors: one
    # candidate run of 24,000 rows, far longer than an LDS tile
    hub = [(0, i) for i in range(1, 6001)]
    fan = [(i, 10_000 + (i * 7 + j) % 3000) for i in range(1, 6001) for j in range(4)]
    chain_tail = [(10_000 + i, 10_000 + i + 1) for i in range(40)]
    e = sorted(set(hub + fan + chain_tail))
    edges = cols(e)
    store, summ = _tc(edges, stream)
    rule0 = rule(1, 1, 0)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule0], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


@pytest.mark.parametrize("sparse_probe_rows", [None, 0])
def test_mixed_run_lengths(sparse_probe_rows, stream):
    # sources with 200..1,500 first-layer successors, each with 4 successors
    # of its own: candidate runs of 800..6,000 rows, on both sides of the
    # tile size and of the largest tile sorted in LDS, packed back to back
    # (sparse_probe_rows=0: every iteration probes FULL by its columns)
    rng = np.random.default_rng(0x7A1)
    fans = rng.permutation(np.repeat([200, 700, 768, 1000, 1023, 1024, 1025, 1100, 1500], 6))
    e = set()
    for src, fan in enumerate(fans):
        mids = 1000 + rng.choice(20_000, size=int(fan), replace=False)
        e.update((src, int(m)) for m in mids)
    for m in range(1000, 21_000):
        e.update((m, 30_000 + int(t)) for t in rng.choice(8000, size=4, replace=False))
    edges = cols(sorted(e))
    store, summ = _tc(edges, stream, sparse_probe_rows=sparse_probe_rows)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


@pytest.mark.parametrize("sparse_probe_rows,spread", [(None, False), (0, False), (None, True)])
def test_medium_candidate_runs(sparse_probe_rows, spread, stream):
    # sources with 1,025..4,097 first-layer successors, each with 4 successors
    # drawn from 6,000 targets (so many candidates repeat): candidate runs of
    # 4,100..16,388 rows, from just past the small LDS tile to just past the
    # largest tile one block sorts; a chain behind some targets adds
    # iterations whose candidates meet FULL. spread: ids scaled by 65,537,
    # so a run's second columns differ over more than 20 bits (too wide for
    # the in-LDS sort keys: such tiles go to the global sort)
    rng = np.random.default_rng(0x3ED)
    fans = rng.permutation(np.repeat([1025, 1500, 2048, 3000, 4095, 4096, 4097], 2))
    e = set()
    for src, fan in enumerate(fans):
        mids = 100 + rng.choice(12_000, size=int(fan), replace=False)
        e.update((src, int(m)) for m in mids)
    for m in range(100, 12_100):
        e.update((m, 20_000 + int(t)) for t in rng.choice(6000, size=4, replace=False))
    e.update((t, t + 1) for t in range(20_000, 20_030))
    if spread:
        e = {(a * 65_537, b * 65_537) for a, b in e}
    edges = cols(sorted(e))
    store, summ = _tc(edges, stream, sparse_probe_rows=sparse_probe_rows)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


@pytest.mark.parametrize("spread", [1, 1 << 17])
def test_long_run_with_tail_at_high_ids(spread, stream):
    # ids above 2^30 (32-bit minors): hub b reaches 5,000 nodes b+1.. in one
    # step, each of those reaches one node b+6000+i, which reaches
    # c + i*spread. The first iteration's candidates are b's run of 5,000
    # (minors agreeing above their low 18 bits: the run bitmap) followed, in
    # the same long tile, by one-row runs whose minors differ over 30 bits
    # with spread 2^17 (tail tiles too wide for 32-bit sort keys) or over 13
    # bits with spread 1 (compressed keys); the second iteration gives b a
    # run of 5,000 over those same minors (the global sort when wide)
    b, c, n = 1 << 30, 1 << 31, 5000
    e = {(b, b + 1 + i) for i in range(n)}
    e |= {(b + 1 + i, b + 6000 + i) for i in range(n)}
    e |= {(b + 6000 + i, c + i * spread) for i in range(n)}
    edges = cols(sorted(e))
    store, summ = _tc(edges, stream)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


@pytest.mark.parametrize("sparse_probe_rows", [None, 0])
def test_huge_segment_probe(sparse_probe_rows, stream):
    # node 0 reaches 600,000 nodes in one step: FULL's segment of first value
    # 0 is longer than a tile streams (2^19 rows), so the candidates of 0 in
    # the next iteration are searched in place; the other first values
    # (the layer-1 nodes, one successor each) are probed through LDS
    n1 = 600_000
    layer1 = np.arange(1, n1 + 1, dtype=np.uint32)
    layer2 = (n1 + 1 + (layer1 % 1000)).astype(np.uint32)
    tail = np.array([n1 + 1, n1 + 1001, n1 + 1002], dtype=np.uint32)  # two more steps
    src = np.concatenate([np.zeros(n1, np.uint32), layer1, tail[:-1]])
    dst = np.concatenate([layer1, layer2, tail[1:]])
    key = np.unique((src.astype(np.uint64) << 32) | dst)
    edges = ((key >> 32).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32))
    store, summ = _tc(edges, stream, sparse_probe_rows=sparse_probe_rows)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


def test_high_degree_join_keys(stream):
    # join keys of degree 250-1,200 next to keys of Poisson(1.5) degree: the
    # join's per-row match counts and the output-balanced emit under skew
    rng = np.random.default_rng(7)
    n = 1500
    deg = rng.poisson(1.5, n)
    deg[[16, 31, 32, 40, 97, 111, 1023]] = [254, 255, 256, 1200, 254, 250, 300]
    deg[rng.choice(np.arange(128, n), 20, replace=False)] = rng.integers(240, 270, 20)
    src = np.repeat(np.arange(n, dtype=np.uint32), deg)
    dst = np.concatenate([rng.choice(n, d, replace=False) for d in deg]).astype(np.uint32)
    key = np.unique((src.astype(np.uint64) << 32) | dst)
    edges = ((key >> 32).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32))
    degrees = np.bincount(edges[0], minlength=n)
    assert degrees[31] == 255 and degrees[97] == 254 and degrees[40] == 1200
    store, summ = _tc(edges, stream)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


def test_join_key_past_narrow_offsets(stream):
    # a key with 2^21 + 8 matches: a count tile of 2,048 rows could then total
    # 2^32 or more, so the join keeps its per-tile offsets as u64
    n = (1 << 21) + 8
    src = np.concatenate([np.zeros(n, np.uint32), np.array([n + 1, n + 2], np.uint32)])
    dst = np.concatenate([np.arange(1, n + 1, dtype=np.uint32), np.array([0, n + 1], np.uint32)])
    key = np.unique((src.astype(np.uint64) << 32) | dst)
    edges = ((key >> 32).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32))
    store, summ = _tc(edges, stream)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [rule(1, 1, 0)], [1])
    full, expect = store.full(R(1)), ref.get(1)
    assert full[0].size == 3 * n + 3
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


def test_ungrouped_projection_plan(stream):
    # output (edge.1, path.0): the distinct's major column is not DELTA's first
    # column, so candidates take the general radix-sort path
    s, d, _ = workload.component_graph(nodes=1 << 12)
    edges = (s, d)
    plan_rule = plans.RecursiveRulePlan(
        target=R(1), delta_input=R(1), right_input=R(0),
        join=plans.BinaryEqualityJoin(1, 0, (plans.JoinProjection(plans.JoinInput.RIGHT, 1),
                                             plans.JoinProjection(plans.JoinInput.LEFT, 0))))
    store = RelationStore(stream)
    store.insert_static(R(0), edges)
    store.insert_recursive(R(1), edges, edges)
    summ = RecursiveExecutor.compile(plans.RecursiveSccPlan((R(1),), (plan_rule,)), store).run(store, 64)
    ref, rsum = _oracle_run({0: edges}, {1: (edges, edges)}, [plan_rule], [1], 64)
    full, expect = store.full(R(1)), ref.get(1)
    assert np.array_equal(full[0], expect[0]) and np.array_equal(full[1], expect[1])
    assert summ.iterations == rsum.iterations and summ.total_new_rows == rsum.total_new_rows


@pytest.mark.parametrize("merge_ratio,segmented,indexed,sparse", [(2, True, True, 1 << 21), (4, False, True, 0),
                                                                  (4, True, False, 0), (4, True, False, 1 << 40),
                                                                  (16, False, False, 1 << 21)])
def test_tuning_does_not_change_results(merge_ratio, segmented, indexed, sparse, expected, stream):
    e = expected["component_graph_16384"]
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store = RelationStore(stream)
    store.set_tuning(merge_ratio=merge_ratio, segmented_distinct=segmented, indexed_anti_join=indexed,
                     sparse_probe_rows=sparse)
    store.insert_static(R(0), (s, d))
    store.insert_recursive(R(1), (s, d), (s, d))
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 10_000)
    full = store.full(R(1))
    bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
    assert np.array_equal(full[0], bfs["closure"][0]) and np.array_equal(full[1], bfs["closure"][1])
    assert summ.iterations == e["max_dist"]
    assert summ.total_new_rows == e["path_rows"] - e["edges"]


def test_program_with_negation_matches_general_path(expected, stream):
    # path = TC(edge), nonlocal = path \ edge: the general CPU path's program
    from sparkalog_amd import program

    e = expected["random_dag_1000_4000"]
    edges = workload.random_dag()
    store = RelationStore(stream)
    store.insert_static(R(0), edges)
    store.insert_recursive(R(1), edges, edges)
    summ = program.run_program(program.transitive_closure_nonlocal(R(1), R(0), R(2)), store)
    g = oracle.general_tc(*edges)
    nl = store.full(R(2))
    assert np.array_equal(nl[0], g["nonlocal"][0]) and np.array_equal(nl[1], g["nonlocal"][1])
    assert summ.derived_rows[2] == e["nonlocal_rows"]
    assert summ.scc[0].iterations == g["iterations"]
    # an empty negative side keeps every row; an empty positive side gives none
    store.insert_static(R(3), (np.empty(0, np.uint32), np.empty(0, np.uint32)))
    assert store.difference(R(4), R(0), R(3)) == edges[0].size
    assert store.difference(R(5), R(3), R(0)) == 0


def _wide_expected(s, d, nodes, labels):
    bfs = oracle.tc_bfs(s, d, nodes, closure=True)
    a, b = bfs["closure"]
    return np.sort((labels[a].astype(np.uint64) << np.uint64(32)) | labels[b].astype(np.uint64)), bfs


@pytest.mark.parametrize("layout", ["auto", "hashed"])
def test_ids_over_the_whole_u32_range(layout, stream):
    # labels spread over [0, 2^32), 0 and 0xFFFFFFFF among them: the join
    # index holds the unique keys through its hash table (the reference
    # indexes any u32 key, storage/src/lib.rs:357-393); exact arrays against
    # BFS and the oracle's semi-naive executor
    nodes = 1 << 12
    s, d, _ = workload.component_graph(nodes=nodes)
    labels = workload.wide_labels(nodes)
    edges = workload.relabel(s, d, labels)
    store = RelationStore(stream)
    store.set_join_index(layout)
    store.insert_static(R(0), edges)
    store.insert_recursive(R(1), edges, edges)
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 10_000)
    exp, bfs = _wide_expected(s, d, nodes, labels)
    full = store.full(R(1))
    assert np.array_equal((full[0].astype(np.uint64) << np.uint64(32)) | full[1], exp)
    assert summ.iterations == bfs["max_dist"]
    assert summ.total_new_rows == exp.size - edges[0].size
    ref = oracle.Store()
    ref.insert_static(0, edges)
    ref.insert_recursive(1, edges, edges)
    rc, ref_summary = ref.run([1], [oracle.rule_from_plan(plans.transitive_closure_scc(R(1), R(0)).rules[0])],
                              10_000, 4)
    assert rc == 0
    o = ref.get(1)
    assert np.array_equal(full[0], o[0]) and np.array_equal(full[1], o[1])
    assert summ.iterations == ref_summary.iterations


def test_ids_over_the_whole_u32_range_2p17(expected, stream):
    # the 2^17-node component graph relabelled into the full u32 range:
    # rows, order-independent digest, sortedness and iterations against the
    # committed BFS fixture
    e = expected["component_graph_131072_wide"]
    s, d, _ = workload.component_graph(nodes=1 << 17)
    edges = workload.relabel(s, d, workload.wide_labels(1 << 17))
    assert workload.pair_digest(*edges) == e["edge_digest"]
    store, summ = _tc(edges, stream)
    digest, sorted_unique, rows = store.digest(R(1))
    assert rows == e["path_rows"] and digest == e["path_digest"] and sorted_unique
    assert summ.iterations == e["max_dist"]
    assert summ.total_new_rows == e["path_rows"] - e["edges"]


@pytest.mark.parametrize("layout", ["dense", "hashed"])
def test_join_index_layouts_agree(layout, expected, stream):
    # the 2^14-node graph with either index layout forced: exact arrays
    e = expected["component_graph_16384"]
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store = RelationStore(stream)
    store.set_join_index(layout)
    store.insert_static(R(0), (s, d))
    store.insert_recursive(R(1), (s, d), (s, d))
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 10_000)
    full = store.full(R(1))
    bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
    assert np.array_equal(full[0], bfs["closure"][0]) and np.array_equal(full[1], bfs["closure"][1])
    assert summ.iterations == e["max_dist"]


def test_wide_ids(stream):
    # node ids spread over [0, 2^28): 64-bit sort keys in the tile sort, no
    # first-column index (ids >= 2^24), the densest join index still allowed
    s, d, _ = workload.component_graph(nodes=1 << 12)
    rng = np.random.default_rng(0x1D5)
    ids = np.sort(rng.choice(1 << 28, size=1 << 12, replace=False)).astype(np.uint32)
    rng.shuffle(ids)
    ws, wd = ids[s], ids[d]
    key = np.unique((ws.astype(np.uint64) << 32) | wd)
    edges = ((key >> 32).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32))
    store, summ = _tc(edges, stream)
    bfs = oracle.tc_bfs(s, d, 1 << 12, closure=True)
    exp = np.unique((ids[bfs["closure"][0]].astype(np.uint64) << 32) | ids[bfs["closure"][1]])
    full = store.full(R(1))
    assert np.array_equal((full[0].astype(np.uint64) << 32) | full[1], exp)
    assert summ.total_new_rows == exp.size - key.size
