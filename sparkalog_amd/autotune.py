"""Executor autotuner: chooses the store's tuning knobs by measurement.

The reference fixes its iteration policies as constants
(crates/sparkalog-recursion/src/lib.rs:24-33); here every knob of the device
executor that leaves results unchanged (sparkalog_hip_store_set_option) is
chosen per workload profile by coordinate descent on a calibration draw:
starting from the built-in defaults, each knob in turn takes every candidate
value while the others stay fixed, and keeps the fastest (median of
`repeats` timed runs after a warm one) when it beats the current choice by
more than `min_gain`; passes repeat until one changes nothing. Every
measured configuration must give the same result signature (rows, digest,
iterations) as the first, so a sweep is also a check that no knob changes
results.

Profiles:
  closure  one large transitive closure (a component graph of 2^20 nodes,
           drawn with a seed other than the benchmark's)
  batched  waves of many small closures (batched and served lanes)

The chosen values are written to a JSON tuning file; RelationStore applies
the file's profile when it is created (`tuning=` argument, else the file
named by SPARKALOG_HIP_TUNING, else the package's tuning_gfx950.json when
present).
"""
from __future__ import annotations

import json
import os
import statistics
import time
from pathlib import Path

import numpy as np

# knob -> candidate values; the first of each is the built-in default
KNOBS: dict[str, tuple[int, ...]] = {
    "distinct_tile": (3072, 1024, 2048, 4096),
    "merge_ratio": (3, 2, 4, 6, 8),
    "sparse_probe_rows": (1 << 22, 1 << 18, 1 << 20, 1 << 24),
    "indexed_anti_join": (0, 1),
    "segmented_distinct": (1, 0),
    "graph": (0, 1),
    "graph_max_rows": (1 << 25, 1 << 22, 1 << 24, 1 << 26),
}
DEFAULTS: dict[str, int] = {k: v[0] for k, v in KNOBS.items()}
PROFILES = ("closure", "batched")
PACKAGE_FILE = Path(__file__).resolve().parent / "tuning_gfx950.json"
ENV_FILE = "SPARKALOG_HIP_TUNING"

# calibration draws (seeds differ from the benchmark workloads')
CALIBRATION_SEED = 0xCA11B
CLOSURE_NODES = 1 << 20
BATCHED_INSTANCES = 512
BATCHED_BUDGET_PAIRS = 3e9


def apply(store, values: dict) -> None:
    """Set every knob of `values` on `store` (unknown names raise)."""
    for name, value in values.items():
        if name not in KNOBS:
            raise KeyError(f"unknown tuning knob {name!r}")
        store.set_option(name, int(value))


def read(path) -> dict:
    data = json.loads(Path(path).read_text())
    if data.get("format") != "sparkalog-hip-tuning/1":
        raise ValueError(f"{path}: not a tuning file")
    for prof in data["profiles"].values():
        for name in prof["values"]:
            if name not in KNOBS:
                raise ValueError(f"{path}: unknown knob {name!r}")
    return data


def default_file() -> Path | None:
    """The tuning file a new store loads: SPARKALOG_HIP_TUNING (empty = none),
    else the package's file when present."""
    env = os.environ.get(ENV_FILE)
    if env is not None:
        return Path(env) if env else None
    return PACKAGE_FILE if PACKAGE_FILE.exists() else None


def profile_values(tuning, profile: str) -> dict:
    """Knob values of `profile` from `tuning`: None (built-in defaults), a
    dict of values, a tuning-file dict, or a path."""
    if tuning is None:
        return {}
    if isinstance(tuning, (str, Path)):
        tuning = read(tuning)
    if "profiles" in tuning:
        prof = tuning["profiles"].get(profile)
        return dict(prof["values"]) if prof else {}
    return dict(tuning)


def coordinate_descent(measure, knobs: dict = KNOBS, start: dict | None = None, min_gain: float = 0.02,
                       max_passes: int = 3, log=None):
    """Minimise measure(values) -> (ms, signature) over `knobs`. Returns
    (best values, best ms, trace). Raises when two configurations give
    different signatures."""
    cur = dict(DEFAULTS if start is None else start)
    cur = {k: cur[k] for k in knobs}
    cache: dict[tuple, float] = {}
    ref_sig = None
    trace = []

    def run(values):
        nonlocal ref_sig
        key = tuple(sorted(values.items()))
        if key in cache:
            return cache[key]
        ms, sig = measure(values)
        if ref_sig is None:
            ref_sig = sig
        elif sig != ref_sig:
            raise AssertionError(f"tuning {values} changed the result: {sig} != {ref_sig}")
        cache[key] = ms
        trace.append({"values": dict(values), "ms": round(ms, 3)})
        if log:
            log(f"  {values} -> {ms:.3f} ms")
        return ms

    best_ms = run(cur)
    for _ in range(max_passes):
        changed = False
        for knob, candidates in knobs.items():
            if knob == "graph_max_rows" and not cur.get("graph", 0):
                continue  # inert without graph mode
            for v in candidates:
                if v == cur[knob]:
                    continue
                trial = dict(cur)
                trial[knob] = v
                ms = run(trial)
                if ms < best_ms * (1.0 - min_gain):
                    cur, best_ms, changed = trial, ms, True
        if not changed:
            break
    return cur, best_ms, trace


def _timed(fn, repeats: int) -> tuple[float, object]:
    import torch

    sig = fn()  # warm: grows workspaces and pools
    times = []
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s = fn()
        torch.cuda.synchronize()
        times.append(1000.0 * (time.perf_counter() - t0))
        if s != sig:
            raise AssertionError("repeated runs disagree")
    return statistics.median(times), sig


def closure_measure(nodes: int = CLOSURE_NODES, seed: int = CALIBRATION_SEED, repeats: int = 3):
    """measure() for the closure profile: TC of a component graph, inputs
    resident in HBM, each configuration on a fresh store."""
    import torch

    from . import plans, workload
    from .execution import HipStream
    from .recursion import RecursiveExecutor, RelationStore

    src, dst, _ = workload.component_graph(seed=seed, nodes=nodes)
    d_src = torch.from_numpy(src.view(np.int32)).cuda()
    d_dst = torch.from_numpy(dst.view(np.int32)).cuda()
    R = plans.RelationId
    plan = plans.transitive_closure_scc(R(1), R(0))

    def measure(values):
        store = RelationStore(HipStream(), tuning=values)

        def once():
            store.clear()
            store.insert_static(R(0), (d_src, d_dst))
            store.insert_recursive(R(1), (d_src, d_dst), (d_src, d_dst))
            summ = RecursiveExecutor.compile(plan, store).run(store, 100_000)
            digest, sorted_unique, rows = store.digest(R(1))
            return rows, digest, sorted_unique, summ.iterations

        try:
            return _timed(once, repeats)
        finally:
            store.close()

    return measure


def batched_measure(instances: int = BATCHED_INSTANCES, seed: int = CALIBRATION_SEED, repeats: int = 3,
                    budget_pairs: float = BATCHED_BUDGET_PAIRS):
    """measure() for the batched profile: `instances` closures of the batched
    workload's size distribution, in waves of the batched benchmark's HBM
    budget."""
    from . import batch, workload
    from .execution import HipStream
    from .recursion import RelationStore

    s, d, offs, counts = workload.batch_graphs(seed=seed, instances=instances)

    def measure(values):
        store = RelationStore(HipStream(), tuning=values)

        def once():
            res = batch.run_batched(s, d, offs, counts, store, budget_pairs=budget_pairs)
            return int(res.rows.sum()), int(np.bitwise_xor.reduce(res.digests)), int(res.iterations.max())

        try:
            return _timed(once, repeats)
        finally:
            store.close()

    return measure


def tune(profiles=PROFILES, log=None, repeats: int = 3) -> dict:
    """Run the coordinate descent for each profile; returns a tuning-file
    dict (format sparkalog-hip-tuning/1)."""
    import torch

    out = {"format": "sparkalog-hip-tuning/1", "device": torch.cuda.get_device_name(0),
           "arch": getattr(torch.cuda.get_device_properties(0), "gcnArchName", ""), "profiles": {}}
    for prof in profiles:
        if log:
            log(f"profile {prof}")
        measure = closure_measure(repeats=repeats) if prof == "closure" else batched_measure(repeats=repeats)
        t0 = time.perf_counter()
        values, ms, trace = coordinate_descent(measure, log=log)
        default_ms = next(t["ms"] for t in trace if t["values"] == DEFAULTS)
        out["profiles"][prof] = {
            "values": values, "ms": round(ms, 3), "default_ms": default_ms,
            "speedup_vs_default": round(default_ms / ms, 4), "configurations": len(trace),
            "seconds": round(time.perf_counter() - t0, 1), "trace": trace,
            "calibration": ({"workload": "component graph", "nodes": CLOSURE_NODES, "seed": CALIBRATION_SEED}
                            if prof == "closure" else
                            {"workload": "batched closures", "instances": BATCHED_INSTANCES,
                             "budget_pairs": BATCHED_BUDGET_PAIRS, "seed": CALIBRATION_SEED}),
        }
    return out


def write(tuning: dict, path) -> None:
    Path(path).write_text(json.dumps(tuning, indent=1) + "\n")
