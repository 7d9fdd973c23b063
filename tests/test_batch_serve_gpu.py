"""Batched and served fixpoints on the GPU.

Batched: many independent TC instances in one semi-naive loop (in waves
under an HBM budget) give, per instance, the same closure rows, digest and
iteration count as brute-force BFS and as the plain one-instance-per-run
path; the full 4,096-instance batch matches committed BFS fixtures.
Served: every request's result is the same alone (plain, sequential) and in
the continuously batched mix, and equals BFS.
"""
from pathlib import Path

import numpy as np
import pytest

import oracle
from sparkalog_amd import batch, plans, serving, workload
from sparkalog_amd.recursion import RelationStore

R = plans.RelationId

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def bfs_per_instance(s, d, offs, counts):
    out = []
    for i in range(counts.size):
        r = oracle.tc_bfs(s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]], int(counts[i]), threads=8)
        out.append((r["pairs"], r["digest"], r["max_dist"]))
    return np.array(out, dtype=np.uint64)


def test_batched_matches_bfs_and_plain(stream):
    s, d, offs, counts = workload.batch_graphs(seed=0x5B, instances=160, n_max=2048)
    ref = bfs_per_instance(s, d, offs, counts)
    store = RelationStore(stream)
    total_est = sum(batch.estimate_closure(int(n), int(e)) for n, e in zip(counts, np.diff(offs.astype(np.int64))))
    res = batch.run_batched(s, d, offs, counts, store, budget_pairs=total_est / 3)
    assert len(res.waves) >= 3
    assert np.array_equal(res.rows, ref[:, 0])
    assert np.array_equal(res.digests, ref[:, 1])
    assert np.array_equal(res.iterations.astype(np.uint64), ref[:, 2])
    plain = batch.run_plain(s, d, offs, counts, store)
    assert np.array_equal(plain.rows, res.rows) and np.array_equal(plain.digests, res.digests)
    assert np.array_equal(plain.iterations, res.iterations)


def test_instance_stats_follow_a_change_to_the_relation(stream):
    # the statistics a run accumulates answer instance_stats only while the
    # relation is unchanged: a pair added to one instance's FULL afterwards
    # shows in that instance's rows and digest, and nowhere else
    from sparkalog_amd.recursion import RecursiveExecutor

    s, d, offs, counts = workload.batch_graphs(seed=0x5C, instances=24, n_max=512)
    ref = bfs_per_instance(s, d, offs, counts)
    bits = batch.local_bits_for(counts)
    members = list(range(counts.size))
    gs, gd = batch._wave_edges(s, d, offs, members, bits)
    store = RelationStore(stream)
    store.insert_static(batch.EDGE, (gs, gd))
    store.insert_recursive(batch.PATH, (gs, gd), (gs, gd))
    store.set_instance_tracking(bits, len(members))
    RecursiveExecutor.compile(batch.TC_PLAN, store).run(store, 100_000)
    rows, digests = store.instance_stats(batch.PATH, bits, len(members))
    assert np.array_equal(rows, ref[:, 0]) and np.array_equal(digests, ref[:, 1])
    first, second = store.full(batch.PATH)
    have = set(zip(first.tolist(), second.tolist()))
    n0 = int(counts[0])
    pair = next((x, z) for x in range(n0) for z in range(n0) if (x, z) not in have)
    store.merge_rows(batch.PATH, 0, (np.array([pair[0]], np.uint32), np.array([pair[1]], np.uint32)))
    rows2, digests2 = store.instance_stats(batch.PATH, bits, len(members))
    assert rows2[0] == rows[0] + 1 and digests2[0] != digests[0]
    assert np.array_equal(rows2[1:], rows[1:]) and np.array_equal(digests2[1:], digests[1:])


def test_full_batch_matches_fixture(stream):
    exp = np.load(GOLDEN / "batch_expected.npz")
    s, d, offs, counts = workload.batch_graphs()
    assert workload.pair_digest(s, d) == int(exp["edge_digest"])
    store = RelationStore(stream)
    res = batch.run_batched(s, d, offs, counts, store, budget_pairs=4.5e9)
    assert int(res.rows.sum()) == int(exp["pairs"].sum())
    assert np.array_equal(res.rows, exp["pairs"])
    assert np.array_equal(res.digests, exp["digest"])
    assert np.array_equal(res.iterations, exp["max_dist"])
    # spot-check the plain path on every 256th instance
    picks = list(range(0, counts.size, 256))
    plain = batch.run_plain(s, d, offs, counts, store, instances=picks)
    assert np.array_equal(plain.rows[picks], exp["pairs"][picks])
    assert np.array_equal(plain.digests[picks], exp["digest"][picks])


def test_served_alone_equals_mix(stream):
    reqs = serving.make_requests(24, rate_per_s=400.0, seed=0x5E, min_log2_edges=8, max_log2_edges=13)
    store = RelationStore(stream)
    mixed, _ = serving.serve_batched(reqs, store, budget_pairs=1e9, max_slots=64)
    alone, _ = serving.serve_plain(reqs, store)
    assert [r.rid for r in mixed] == list(range(24))
    for m, a, req in zip(mixed, alone, reqs):
        assert (m.rows, m.digest) == (a.rows, a.digest), req.rid
        bfs = oracle.tc_bfs(req.src, req.dst, req.nodes)
        assert (m.rows, m.digest) == (bfs["pairs"], bfs["digest"]), req.rid
    # requests overlapped: some were admitted before an earlier one finished
    assert any(mixed[i + 1].admitted < mixed[i].done for i in range(len(mixed) - 1))


def test_served_lanes_equal_alone(stream):
    # three size-class lanes, each on its own stream and host thread
    import torch
    from sparkalog_amd.execution import HipStream

    reqs = serving.make_requests(30, rate_per_s=600.0, seed=0x5F, min_log2_edges=8, max_log2_edges=14)
    stores = [RelationStore(HipStream(torch.cuda.Stream())) for _ in range(3)]
    mixed, _ = serving.serve_lanes(reqs, stores, bounds=[2e4, 2e5], budget_pairs=1e9, max_slots=64)
    alone, _ = serving.serve_plain(reqs, RelationStore(stream))
    assert [r.rid for r in mixed] == list(range(30))
    for m, a, req in zip(mixed, alone, reqs):
        assert (m.rows, m.digest) == (a.rows, a.digest), req.rid


def test_served_requests_up_to_2p20_edges(stream):
    # the served stream's full edge range: eight requests with edges
    # log-uniform on [2^8, 2^20] (seed 0x5C21 draws 424 to 1,011,046 edges),
    # arriving faster than they are served, in three lanes: every result
    # alone = in the mix = BFS
    import torch
    from sparkalog_amd.execution import HipStream

    reqs = serving.make_requests(8, rate_per_s=40.0, seed=0x5C21, min_log2_edges=8, max_log2_edges=20)
    assert max(r.src.size for r in reqs) > 1_000_000
    stores = [RelationStore(HipStream(torch.cuda.Stream())) for _ in range(3)]
    mixed, _ = serving.serve_lanes(reqs, stores, bounds=[1e6, 3e7], budget_pairs=6e9, max_slots=64)
    alone, _ = serving.serve_plain(reqs, RelationStore(stream))
    assert [r.rid for r in mixed] == list(range(len(reqs)))
    for m, a, req in zip(mixed, alone, reqs):
        bfs = oracle.tc_bfs(req.src, req.dst, req.nodes, threads=8)
        assert (m.rows, m.digest) == (a.rows, a.digest) == (bfs["pairs"], bfs["digest"]), req.rid


def test_served_lanes_by_edge_count_with_adaptive_admission(stream):
    # routed by edge count (the bench's lanes), admission estimates learned
    # from finished requests: many requests share a batch, results unchanged
    import torch
    from sparkalog_amd.execution import HipStream

    reqs = serving.make_requests(40, rate_per_s=2000.0, seed=0x5E, min_log2_edges=8, max_log2_edges=16)
    stores = [RelationStore(HipStream(torch.cuda.Stream())) for _ in range(3)]
    stats = []
    mixed, _ = serving.serve_lanes(reqs, stores, bounds=[2048, 16384], budget_pairs=3e9, lane_slots=[256, 16, 0],
                                   stats=stats, route="edges")
    alone, _ = serving.serve_plain(reqs, RelationStore(stream))
    assert [r.rid for r in mixed] == list(range(len(reqs)))
    for m, a, req in zip(mixed, alone, reqs):
        bfs = oracle.tc_bfs(req.src, req.dst, req.nodes, threads=8)
        assert (m.rows, m.digest) == (a.rows, a.digest) == (bfs["pairs"], bfs["digest"]), req.rid
    assert sum(st["requests"] for st in stats) == len(reqs)
    assert max(st.get("max_active", 0) for st in stats) > 1  # requests did share a batch
    assert stats[2].get("plain")  # the largest class ran as a plain lane


def test_remove_ranges_matches_filter(stream):
    rng = np.random.default_rng(0x77)
    a = rng.integers(0, 4000, 20000).astype(np.uint32)
    b = rng.integers(0, 4000, 20000).astype(np.uint32)
    key = np.unique((a.astype(np.uint64) << 32) | b)
    cols = ((key >> 32).astype(np.uint32), (key & 0xFFFFFFFF).astype(np.uint32))
    store = RelationStore(stream)
    store.insert_recursive(R(1), cols, cols)
    bounds = np.array([0, 10, 500, 777, 1000, 1001, 3990, 5000], dtype=np.uint32)
    store.remove_ranges(R(1), bounds)
    x = cols[0]
    keep = np.ones(x.size, bool)
    for lo, hi in bounds.reshape(-1, 2):
        keep &= ~((x >= lo) & (x < hi))
    full = store.full(R(1))
    assert np.array_equal(full[0], cols[0][keep]) and np.array_equal(full[1], cols[1][keep])
    assert store.rows(R(1), 1) == int(keep.sum())
