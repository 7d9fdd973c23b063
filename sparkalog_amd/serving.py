"""Served fixpoint queries with continuous batching.

Transitive-closure requests arrive as a Poisson stream. The server keeps one
device-resident relation store holding every in-flight request as an
independent instance (request slot s owns global ids [s << b, (s+1) << b)),
and runs one semi-naive step at a time over all of them:

  * between steps, requests that have arrived are admitted while a slot and
    the HBM budget (estimated closure pairs) allow: their edges are merged
    into the edge relation and their seeds join FULL (as a run of their own)
    and DELTA;
  * after a step, every instance whose DELTA is empty has converged: its
    rows and digest are read from FULL; its rows are removed later, together
    with those of other converged instances, and only then is its slot
    reused.

So requests that arrive while others are on the GPU join the running batch
at the next step, and their work overlaps on the device. Latency is measured
from arrival to the moment its result has been read.

Requests can be split into size-class lanes (serve_lanes), each a
continuous batch of its own on its own stream and host thread, so that small
requests do not step at the pace of the largest instances.

The plain baseline serves the same stream one request at a time, in arrival
order, each with its own executor run.
"""
from __future__ import annotations

import math
import threading
import time
from dataclasses import dataclass

import numpy as np

from . import batch, plans, workload
from .recursion import RecursiveExecutor, RelationStore

R = plans.RelationId
EDGE, PATH = R(0), R(1)


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the stream starts
    src: np.ndarray
    dst: np.ndarray
    nodes: int


@dataclass
class Served:
    rid: int
    arrival: float
    admitted: float
    done: float
    rows: int
    digest: int

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def make_requests(count: int, rate_per_s: float, seed: int = workload.C5_SEED, min_log2_edges: int = 8,
                  max_log2_edges: int = 20) -> list[Request]:
    """Poisson arrivals; edge counts log-uniform in [2^min, 2^max]; each
    request is a component graph (Zipf(1.5) sizes on [8, 4096], Poisson(2)
    out-degree) with nodes = edges / 2 and seed = seed + request id."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count)
    log_edges = rng.uniform(min_log2_edges, max_log2_edges, size=count)
    out, t = [], 0.0
    for i in range(count):
        t += float(gaps[i])
        nodes = max(16, int(2 ** log_edges[i] / 2))
        s, d, _ = workload.component_graph(seed=seed + i, nodes=nodes)
        out.append(Request(i, t, s, d, nodes))
    return out


def percentile(values, q: float) -> float:
    v = sorted(values)
    if not v:
        return float("nan")
    return v[min(len(v) - 1, int(math.ceil(q / 100.0 * len(v))) - 1)]


def summarize(results: list[Served], elapsed: float) -> dict:
    lat = [r.latency * 1000.0 for r in results]
    tuples = sum(r.rows for r in results)
    return {"requests": len(results), "p50_ms": round(percentile(lat, 50), 3), "p99_ms": round(percentile(lat, 99), 3),
            "mean_ms": round(float(np.mean(lat)) if lat else float("nan"), 3), "seconds": round(elapsed, 4),
            "tuples": int(tuples), "tuples_per_s": round(tuples / elapsed, 1) if elapsed > 0 else None}


def _wait_until(t0: float, t: float) -> None:
    delay = t0 + t - time.perf_counter()
    if delay > 0:
        time.sleep(delay)


def _plain_loop(requests: list[Request], store: RelationStore, t0: float, results: list[Served],
                stats: dict | None = None) -> None:
    """One executor run per request, in arrival order, on `store`."""
    plan = plans.transitive_closure_scc(PATH, EDGE)
    busy = 0.0
    for req in requests:
        _wait_until(t0, req.arrival)
        start = time.perf_counter() - t0
        store.clear()
        store.insert_static(EDGE, (req.src, req.dst))
        store.insert_recursive(PATH, (req.src, req.dst), (req.src, req.dst))
        RecursiveExecutor.compile(plan, store).run(store, 1_000_000)
        digest, _, rows = store.digest(PATH)
        done = time.perf_counter() - t0
        busy += done - start
        results.append(Served(req.rid, req.arrival, start, done, rows, digest))
    if stats is not None:
        stats.update({"plain": True, "busy_s": round(busy, 4)})


def serve_plain(requests: list[Request], store: RelationStore) -> tuple[list[Served], float]:
    results: list[Served] = []
    t0 = time.perf_counter()
    _plain_loop(requests, store, t0, results)
    return results, time.perf_counter() - t0


class ClosureEstimator:
    """HBM budgeting estimate of a request's closure: the random-digraph
    estimate batch.estimate_closure (its giant component squared: safe, and
    far above the closure of graphs made of many small components) scaled
    by what finished requests of about the same size showed. Requests are
    bucketed by log2 of their node count; a bucket's factor is the largest
    observed closure / estimate in it, and a request takes `margin` times
    the largest factor of its bucket and the two beside it, at most 1 (the
    plain estimate when none of them has finished a request yet)."""

    def __init__(self, margin: float = 2.0):
        self.margin = margin
        self.factor: dict[int, float] = {}

    @staticmethod
    def _bucket(req: Request) -> int:
        return max(int(req.nodes), 1).bit_length()

    def __call__(self, req: Request) -> float:
        prior = batch.estimate_closure(req.nodes, req.src.size)
        b = self._bucket(req)
        seen = [self.factor[k] for k in (b - 1, b, b + 1) if k in self.factor]
        if not seen:
            return prior
        return prior * min(1.0, self.margin * max(seen))

    def observe(self, req: Request, rows: int) -> None:
        prior = batch.estimate_closure(req.nodes, req.src.size)
        if prior > 0:
            b = self._bucket(req)
            self.factor[b] = max(self.factor.get(b, 0.0), rows / prior)


def _lane(requests: list[Request], store: RelationStore, t0: float, budget_pairs: float, max_slots: int,
          bits: int, results: list[Served], stats: dict | None = None,
          estimator: ClosureEstimator | None = None) -> None:
    """One continuous-batching loop over `requests` (arrival order) on
    `store`'s stream; appends to `results`. `stats`, when given, receives
    the loop's step, admission, retirement and removal counts and seconds."""
    st = {"steps": 0, "step_s": 0.0, "admissions": 0, "admit_s": 0.0, "removals": 0, "remove_s": 0.0,
          "retire_s": 0.0, "idle_s": 0.0, "max_active": 0}
    estimate = estimator or ClosureEstimator()
    if (max_slots << bits) > (1 << 28):
        max_slots = (1 << 28) >> bits
    store.clear()
    empty = (np.empty(0, np.uint32), np.empty(0, np.uint32))
    store.insert_static(EDGE, empty)
    store.insert_recursive(PATH, empty, empty)
    ex = RecursiveExecutor.compile(plans.transitive_closure_scc(PATH, EDGE), store)
    # slot s is instance s: each step reads back every slot's DELTA rows with
    # its own counters, so convergence needs no extra synchronization
    store.set_instance_tracking(bits, max_slots)
    ex.begin(store)
    free = list(range(max_slots))[::-1]
    active = {}  # slot -> (request, admitted time, estimate)
    dead: list[int] = []  # converged slots whose rows are still stored
    dead_pairs = 0
    pending = list(requests)
    used = 0.0
    while pending or active:
        now = time.perf_counter() - t0
        # converged instances' rows are removed in batches: all dead slots in
        # one pass over each relation, when slots run out or their rows grow
        if dead and ((not free and pending) or dead_pairs > budget_pairs / 4):
            tr = time.perf_counter()
            bounds = np.array([[s << bits, (s + 1) << bits] for s in sorted(dead)], dtype=np.uint32).reshape(-1)
            store.remove_ranges(PATH, bounds)
            store.remove_ranges(EDGE, bounds)
            free.extend(dead)
            dead.clear()
            dead_pairs = 0
            st["removals"] += 1
            st["remove_s"] += time.perf_counter() - tr
        # admission between steps, in arrival order: every request that fits
        # joins at once (one merge per relation part for all of them; slots
        # ascend with arrival, so their rows concatenate in sorted order)
        admitted = []
        while pending and pending[0].arrival <= now and free:
            req = pending[0]
            est = estimate(req)
            if (active or admitted) and used + est > budget_pairs:
                break
            pending.pop(0)
            admitted.append((free.pop(), req, est))
            used += est
        if admitted:
            ta = time.perf_counter()
            admitted.sort(key=lambda a: a[0])
            gs = np.concatenate([req.src + np.uint32(slot << bits) for slot, req, _ in admitted])
            gd = np.concatenate([req.dst + np.uint32(slot << bits) for slot, req, _ in admitted])
            store.merge_rows(EDGE, 0, (gs, gd))
            store.merge_rows(PATH, 0, (gs, gd))
            store.merge_rows(PATH, 1, (gs, gd))
            for slot, req, est in admitted:
                active[slot] = (req, now, est)
            st["admissions"] += 1
            st["admit_s"] += time.perf_counter() - ta
        if not active:
            ti = time.perf_counter()
            _wait_until(t0, pending[0].arrival)
            st["idle_s"] += time.perf_counter() - ti
            continue
        st["max_active"] = max(st["max_active"], len(active))
        ts = time.perf_counter()
        ex.step(store)
        st["steps"] += 1
        st["step_s"] += time.perf_counter() - ts
        # retire converged instances (the step's per-instance DELTA rows)
        tr = time.perf_counter()
        left = store.instance_delta_rows(max_slots)
        for slot in list(active):
            if left[slot]:
                continue
            lo, hi = slot << bits, (slot + 1) << bits
            req, admitted_at, est = active.pop(slot)
            rows, digest = store.range_stats(PATH, 0, lo, hi, lo)
            results.append(Served(req.rid, req.arrival, admitted_at, time.perf_counter() - t0, rows, digest))
            estimate.observe(req, rows)
            # its rows stay until the next batched removal: with an empty
            # DELTA they take part in no join, and no new instance uses the
            # slot before they are gone
            used -= est
            dead.append(slot)
            dead_pairs += rows
        st["retire_s"] += time.perf_counter() - tr
    ex.end(store)
    if stats is not None:
        stats.update({k: round(v, 4) if isinstance(v, float) else v for k, v in st.items()})


def _instance_bits(requests: list[Request]) -> int:
    return max(1, max(r.nodes for r in requests) - 1).bit_length()


def serve_batched(requests: list[Request], store: RelationStore, budget_pairs: float = 6e9,
                  max_slots: int = 256) -> tuple[list[Served], float]:
    """Every request in one continuous batch on one stream."""
    results: list[Served] = []
    t0 = time.perf_counter()
    _lane(requests, store, t0, budget_pairs, max_slots, _instance_bits(requests), results)
    results.sort(key=lambda r: r.rid)
    return results, time.perf_counter() - t0


def serve_lanes(requests: list[Request], stores: list[RelationStore], bounds: list[float],
                budget_pairs: float = 6e9, max_slots: int = 256,
                lane_slots: list[int] | None = None, stats: list | None = None,
                route: str = "closure") -> tuple[list[Served], float]:
    """Continuous batching in size classes: request r goes to the first lane
    whose bound exceeds its size - its estimated closure (route "closure")
    or its edge count (route "edges") - and the last lane takes the rest.
    Every request of a component graph takes about as many steps as its
    diameter, and a batch step costs about as much as its largest instances,
    so small requests get lanes of their own; each lane batches on its own
    store and stream, driven by its own host thread (the native calls release
    the interpreter lock), and the HBM budget is split evenly. lane_slots
    caps the instances in flight per lane (default max_slots each); a cap of
    1 serves that lane's requests one at a time, in arrival order, still
    through the batching loop, and a cap of 0 makes the lane a plain loop
    (one executor run per request, as serve_plain)."""
    if len(stores) != len(bounds) + 1:
        raise ValueError("serve_lanes: one store per lane (len(bounds) + 1)")
    lanes: list[list[Request]] = [[] for _ in stores]
    if route not in ("closure", "edges"):
        raise ValueError(f"serve_lanes: unknown route {route!r}")
    for req in requests:
        size = batch.estimate_closure(req.nodes, req.src.size) if route == "closure" else req.src.size
        lane = next((i for i, b in enumerate(bounds) if size < b), len(bounds))
        lanes[lane].append(req)
    bits = _instance_bits(requests)
    results: list[list[Served]] = [[] for _ in stores]
    errors: list[BaseException] = []
    lane_stats = [{"requests": len(lanes[i])} for i in range(len(stores))]
    t0 = time.perf_counter()

    def run(i: int) -> None:
        try:
            if lanes[i]:
                slots = lane_slots[i] if lane_slots else max_slots
                if slots == 0:
                    _plain_loop(lanes[i], stores[i], t0, results[i], lane_stats[i])
                else:
                    _lane(lanes[i], stores[i], t0, budget_pairs / len(stores), slots, bits, results[i],
                          lane_stats[i])
        except BaseException as e:  # re-raised on the calling thread
            errors.append(e)

    threads = [threading.Thread(target=run, args=(i,)) for i in range(len(stores))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    if errors:
        raise errors[0]
    if stats is not None:
        stats.extend(lane_stats)
    merged = sorted((r for lane in results for r in lane), key=lambda r: r.rid)
    return merged, time.perf_counter() - t0
