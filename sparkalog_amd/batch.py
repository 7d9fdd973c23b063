"""Batched fixpoint: thousands of independent transitive-closure instances of
uneven size run as one semi-naive loop.

Instance i's local node v becomes global id (slot << b) | v, b = bits of the
largest local id in its wave, so the batch is the disjoint union of the instances and its
closure is the union of their closures. Sort keys of the union are
(slot, x, y) packed into one u64; the executor's segmented distinct and
tiered FULL apply unchanged, and per-instance results come from one
segmented reduction over FULL. When the estimated closure of the whole batch
does not fit the HBM budget, instances run in waves that do.

The plain path runs the same executor once per instance.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import plans
from .recursion import RecursiveExecutor, RelationStore

R = plans.RelationId
EDGE, PATH = R(0), R(1)
TC_PLAN = plans.transitive_closure_scc(PATH, EDGE)


def giant_fraction(mean_degree: float) -> float:
    """Fraction of nodes in the giant component of a random digraph with the
    given mean out-degree: the root of g = 1 - exp(-lambda g)."""
    if mean_degree <= 1.0:
        return 0.0
    g = 0.9
    for _ in range(100):
        g = 1.0 - math.exp(-mean_degree * g)
    return g


def estimate_closure(nodes: int, edges: int) -> float:
    """Closure-size estimate used only for HBM budgeting (with margin)."""
    if nodes == 0 or edges == 0:
        return 0.0
    g = giant_fraction(edges / nodes)
    return min(float(nodes) * nodes, 1.4 * (g * nodes) ** 2 + 8.0 * nodes)


@dataclass
class BatchResult:
    rows: np.ndarray
    digests: np.ndarray
    iterations: np.ndarray
    seconds: float
    waves: list = field(default_factory=list)


def local_bits_for(node_counts) -> int:
    return max(1, int(np.max(node_counts) - 1).bit_length()) if len(node_counts) else 1


def estimated_iterations(nodes: int, edges: int) -> float:
    """Rough iteration count of an instance's closure (its longest shortest
    path grows as log(n) / log(mean out-degree) in a random digraph); only
    used to order instances into waves."""
    if nodes <= 1 or edges == 0:
        return 0.0
    return math.log(nodes) / math.log(max(edges / nodes, 1.05))


def diameter_order(node_counts, edge_counts) -> list[int]:
    """Instances by estimated iteration count, longest first: a wave runs as
    many iterations as its longest member, so grouping long members together
    saves the other waves their tail of nearly empty iterations."""
    est = [estimated_iterations(int(n), int(e)) for n, e in zip(node_counts, edge_counts)]
    return sorted(range(len(est)), key=lambda i: -est[i])


def plan_waves(node_counts, edge_counts, budget_pairs: float, max_members: int | None = None,
               order=None) -> list[list[int]]:
    """Consecutive instances (in `order`, default index order) per wave while
    their estimated closures fit the budget, and at most max_members
    (default: as many as global ids slot << bits | v fit in u32, bits =
    local_bits_for(node_counts))."""
    if max_members is None:
        max_members = 1 << (32 - local_bits_for(node_counts)) if len(node_counts) else 1
    waves, current, used = [], [], 0.0
    for i in (range(len(node_counts)) if order is None else order):
        n, e = node_counts[i], edge_counts[i]
        est = estimate_closure(int(n), int(e))
        if current and (used + est > budget_pairs or len(current) >= max_members):
            waves.append(current)
            current, used = [], 0.0
        current.append(i)
        used += est
    if current:
        waves.append(current)
    return waves


def _wave_edges(src, dst, offsets, members, bits):
    parts_s, parts_d = [], []
    for slot, i in enumerate(members):
        a, b = offsets[i], offsets[i + 1]
        base = np.uint32(slot << bits)
        parts_s.append(src[a:b] + base)
        parts_d.append(dst[a:b] + base)
    return (np.concatenate(parts_s) if parts_s else np.empty(0, np.uint32),
            np.concatenate(parts_d) if parts_d else np.empty(0, np.uint32))


def run_batched(src, dst, offsets, node_counts, store: RelationStore, budget_pairs: float = 3e9,
                max_iterations: int = 100_000) -> BatchResult:
    n = len(node_counts)
    edge_counts = np.diff(np.asarray(offsets, dtype=np.int64))
    rows = np.zeros(n, np.uint64)
    digests = np.zeros(n, np.uint64)
    iterations = np.zeros(n, np.uint32)
    waves = plan_waves(node_counts, edge_counts, budget_pairs, order=diameter_order(node_counts, edge_counts))
    t0 = time.perf_counter()
    log = []
    for members in waves:
        tw = time.perf_counter()
        # a wave's local ids take the bits of its largest member, so a wave
        # of small instances has a small id space (and small indexes)
        wbits = local_bits_for(np.asarray(node_counts)[members])
        gs, gd = _wave_edges(src, dst, offsets, members, wbits)
        store.clear()
        store.insert_static(EDGE, (gs, gd))
        store.insert_recursive(PATH, (gs, gd), (gs, gd))
        store.set_instance_tracking(wbits, len(members))
        summary = RecursiveExecutor.compile(TC_PLAN, store).run(store, max_iterations)
        r, d = store.instance_stats(PATH, wbits, len(members))
        last = store.instance_iterations(len(members))
        idx = np.asarray(members)
        rows[idx], digests[idx] = r, d
        iterations[idx] = np.where(edge_counts[idx] > 0, last + 1, 0)
        log.append({"instances": len(members), "tuples": int(r.sum()), "iterations": summary.iterations,
                    "device_ms": round(summary.device_ms, 3), "wall_ms": round(1000 * (time.perf_counter() - tw), 3)})
    store.set_instance_tracking(0, 0)
    return BatchResult(rows, digests, iterations, time.perf_counter() - t0, log)


def run_plain(src, dst, offsets, node_counts, store: RelationStore, max_iterations: int = 100_000,
              instances=None) -> BatchResult:
    """One executor run per instance (the one-instance-per-launch baseline)."""
    n = len(node_counts)
    rows = np.zeros(n, np.uint64)
    digests = np.zeros(n, np.uint64)
    iterations = np.zeros(n, np.uint32)
    todo = range(n) if instances is None else instances
    t0 = time.perf_counter()
    for i in todo:
        a, b = offsets[i], offsets[i + 1]
        e = (src[a:b], dst[a:b])
        store.clear()
        store.insert_static(EDGE, e)
        store.insert_recursive(PATH, e, e)
        summary = RecursiveExecutor.compile(TC_PLAN, store).run(store, max_iterations)
        d, _, r = store.digest(PATH)
        rows[i], digests[i], iterations[i] = r, d, summary.iterations
    return BatchResult(rows, digests, iterations, time.perf_counter() - t0)
