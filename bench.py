"""Benchmark: single-instance transitive closure on one MI355X.

One step = one full semi-naive fixpoint of
    path(x,y) :- edge(x,y).   path(x,z) :- path(x,y), edge(y,z).
over the component graph of 2^20 u32 nodes (component sizes Zipf(1.5) on
[8, 4096], Poisson(2.0) out-degree, ids permuted; ~2.09M edges), run by the
device-resident executor. The edge columns are resident in HBM before the
timed region; every step re-seeds FULL = DELTA = edge on the device and runs
to the fixpoint (33 iterations, 1,034,987,993 result tuples).

Prints one JSON line: tuples/sec (whole job), HBM GB/s of algorithmic bytes,
the roofline of the dominant operator (per-launch HIP events on the executor's
stream, recorded in as many further steps as are timed, since the events add
about 3% to a fixpoint), and the CPU baseline (the oracle's parallel
semi-naive executor on a bounded sample, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
PROFILE_SUMMARY = ROOT / "profiles" / "tc_component_graph_2p20_traffic.json"


def parse_args():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--nodes", type=int, default=1 << 20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-batched", action="store_true", help="tc: leave out the batched configuration")
    p.add_argument("--no-served", action="store_true", help="tc: leave out the served configuration")
    p.add_argument("--bench-plain-stride", type=int, default=8,
                   help="tc: the batched summary times the plain path on every k-th instance")
    p.add_argument("--merge-ratio", type=int, default=None,
                   help="tc: FULL run-merge ratio (default: the tuning file's; results are identical)")
    p.add_argument("--host-io", action="store_true",
                   help="tc: also time one fixpoint from host arrays to the closure back in host memory (PCIe)")
    p.add_argument("--indexed-anti-join", type=int, default=None,
                   help="tc: anti-join FULL through its first-column index, 0/1 (default: the tuning file's)")
    p.add_argument("--sparse-probe-rows", type=int, default=None,
                   help="tc: candidate count up to which FULL is probed one candidate per thread "
                        "(default: the tuning file's)")
    p.add_argument("--tuning", default="file",
                   help="tuning file applied to every store ('file' = the package's or SPARKALOG_HIP_TUNING, "
                        "'none' = built-in defaults, or a path)")
    p.add_argument("--tuning-out", default="gpurun_out/tuning_gfx950.json", help="tune mode: file written")
    p.add_argument("--tune-profiles", default="closure,batched", help="tune mode: profiles tuned")
    p.add_argument("--no-default-compare", action="store_true",
                   help="tc: skip timing the built-in default tuning beside the tuned one")
    p.add_argument("--cpu-sample-nodes", type=int, default=1 << 16,
                   help="CPU baseline: nodes of the parallel oracle's component-graph draw")
    p.add_argument("--cpu-serial-nodes", type=int, default=1 << 14,
                   help="CPU baseline: nodes of the serial oracle's component-graph draw")
    p.add_argument("--crossover-ops", default="filter,join,distinct,anti_join,union",
                   help="crossover mode: operators")
    p.add_argument("--crossover-min-log2", type=int, default=8, help="crossover mode: smallest log2 row count")
    p.add_argument("--crossover-max-log2", type=int, default=24, help="crossover mode: largest log2 row count")
    p.add_argument("--chain", type=int, default=2000, help="graph mode: nodes of the chain whose closure is timed")
    p.add_argument("--mode", choices=["tc", "ops", "batched", "served", "copy", "crossover", "tune", "graph"],
                   default="tc",
                   help="tc: the headline fixpoint; ops: operator microbench at 2^k rows; "
                        "batched: 4,096 instances in one loop vs one run each; served: Poisson stream, "
                        "continuous batching vs sequential; copy: plain device copy of --copy-bytes "
                        "(the known-size reference for HBM byte counters); tune: run the autotuner and "
                        "write --tuning-out")
    p.add_argument("--copy-bytes", type=int, default=1 << 30, help="copy mode: bytes copied per launch")
    p.add_argument("--budget-pairs", type=float, default=3e9, help="batched/served: HBM budget in closure pairs")
    p.add_argument("--requests", type=int, default=200, help="served: stream length")
    p.add_argument("--utilization", default="0.7,1.0,1.5",
                   help="served mode: offered loads of the plain path (comma list)")
    p.add_argument("--bench-utilization", default="0.7,1.0,1.5",
                   help="tc: offered loads of the served summary in the bench line")
    p.add_argument("--plain-stride", type=int, default=1, help="batched: time the plain path on every k-th instance")
    p.add_argument("--lane-route", default="edges", choices=["edges", "closure"],
                   help="served: route requests to lanes by edge count or by estimated closure")
    p.add_argument("--lane-bounds", default="4096,65536",
                   help="served: closure-size bounds (pairs) between continuous-batching lanes")
    p.add_argument("--lane-slots", default="256,16,0",
                   help="served: instances in flight per lane (comma list; empty = 256 each; 0 = a plain lane, "
                        "one executor run per request)")
    p.add_argument("--graph", type=int, default=-1,
                   help="batched/served: replay single-rule iterations from captured HIP graphs (1) or not (0); "
                        "-1 = the tuning file's choice")
    p.add_argument("--sizes", default="24,26,28", help="ops mode: log2 row counts")
    p.add_argument("--ops", default="distinct,union,anti_join,join",
                   help="ops mode: operators (distinct, union, anti_join, join, filter)")
    p.add_argument("--samples", type=int, default=5, help="ops mode: timed launches per case")
    p.add_argument("--cpu-max-log2", type=int, default=24,
                   help="ops mode: largest case checked against (and timed with) the oracle on this host; "
                        "larger ones are checked against the committed oracle fixtures")
    return p.parse_args()


def _tuning_arg(args):
    """RelationStore's `tuning` argument from --tuning."""
    return {"file": "file", "none": None}.get(args.tuning, args.tuning)


def _graph_arg(args):
    return None if args.graph < 0 else bool(args.graph)


def _cpu_threads() -> int:
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def oracle_operator(name, host, threads):
    """CPU leg of the operator microbench: the oracle's serial (threads = 1)
    or parallel (Rayon-style chunks) operator on the case's host inputs. Its
    output is also what the GPU result must equal before any timing is kept
    (distinct-crossover.rs:227-234)."""
    import oracle

    if name == "distinct":
        return oracle.distinct(*host["input"], threads=threads)
    if name == "union":
        return oracle.union(host["left"], host["right"], threads=threads)
    if name == "anti_join":
        return oracle.anti_join(host["left"], host["right"], threads=threads)
    if name == "filter":
        from sparkalog_amd.opbench import Filter
        return oracle.filter_rows(host["values"], Filter.OP_GE, Filter.OPERAND, threads=threads)
    return oracle.join(host["left"], host["right"], 1, 0, ((0, 0), (1, 1)), threads=threads)


def run_ops(args) -> None:
    """Operator drop-in microbench: one JSON line per (operator, rows, keys).
    Every case is checked before it is timed: against the oracle run on this
    host (which is also timed, as the case's CPU baseline) up to
    2^--cpu-max-log2 rows, and above that against the digests the oracle
    wrote into tests/golden/op_expected.json."""
    import ctypes

    import numpy as np
    import torch

    from sparkalog_amd import opbench, workload

    torch.cuda.set_device(0)
    threads = _cpu_threads()
    fixtures_path = ROOT / "tests" / "golden" / "op_expected.json"
    fixtures = json.loads(fixtures_path.read_text()) if fixtures_path.exists() else {}
    modes = {"uniform": workload.KEY_UNIFORM, "zipf1.1": workload.KEY_ZIPF}
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for name in args.ops.split(","):
        for k in [int(x) for x in args.sizes.split(",")]:
            # the filter's input is i % 97 (no key distribution); it is checked
            # against the oracle at every size
            for mode_name, mode in (modes.items() if name != "filter" else [("i%97", workload.KEY_UNIFORM)]):
                host = opbench.host_inputs(name, 1 << k, mode)
                op = opbench.make(name, 1 << k, mode, host=host)
                op.launch(stream)
                torch.cuda.synchronize()
                got = op.result()
                check = {}
                if name == "filter":
                    exp = oracle_operator(name, host, threads)
                    check = {"checked_against": "oracle"}
                    ok = np.array_equal(got, exp)
                    del exp
                elif k <= args.cpu_max_log2:
                    t0 = time.perf_counter()
                    exp = oracle_operator(name, host, threads)
                    check = {"checked_against": "oracle", "cpu_parallel_ms": round(1000 * (time.perf_counter() - t0), 3),
                             "cpu_threads": threads}
                    ok = got[0].size == exp[0].size and np.array_equal(got[0], exp[0]) and \
                        np.array_equal(got[1], exp[1])
                    del exp
                else:
                    e = fixtures.get(f"{name}_2p{k}_{mode_name}")
                    if e is None:
                        raise SystemExit(f"ops: no fixture for {name} 2^{k} {mode_name}; run "
                                         "tests/golden/make_op_fixtures.py")
                    ok = got[0].size == e["out_rows"] and workload.pair_digest(*got) == e["digest"]
                    if name == "join":
                        ok = ok and workload.ordered_pair_digest(*got) == e["ordered_digest"]
                    else:
                        key = (got[0].astype(np.uint64) << np.uint64(32)) | got[1]
                        ok = ok and bool(np.all(key[1:] > key[:-1]))
                    check = {"checked_against": "oracle fixture (rows, digests, order)"}
                del got
                if not ok:
                    print(json.dumps({"op": name, "rows": 1 << k, "keys": mode_name, "error": "result differs"}),
                          flush=True)
                    raise SystemExit(1)
                r = opbench.time_op(op, args.samples)
                r.update({"keys": mode_name, "frac_of_hbm_peak": round(r["gbs"] / HBM_PEAK_GBS, 4), "checked": True})
                r.update(check)
                print(json.dumps(r), flush=True)
                del op, host
                torch.cuda.empty_cache()


def batched_summary(budget_pairs: float, plain_stride: int, full_log: bool = False, graph=None,
                    tuning="file") -> dict:
    """Batched fixpoints: 4,096 instances (nodes log-uniform on [16, 16384],
    Poisson out-degree mean uniform on [1.2, 3.0]) in one semi-naive loop per
    HBM wave, against one executor run per instance (every plain_stride-th
    instance timed, the total scaled from them)."""
    import numpy as np
    import torch

    from sparkalog_amd import batch, workload
    from sparkalog_amd.execution import HipStream
    from sparkalog_amd.recursion import RelationStore

    torch.cuda.set_device(0)
    s, d, offs, counts = workload.batch_graphs()
    store = RelationStore(HipStream(), tuning=tuning, profile="batched")
    if graph is not None:
        store.set_graph(graph)
    graph = bool(store.tuning.get("graph", 0)) if graph is None else graph
    # a cold pass first: its waves grow the store's workspaces and pools
    # (multi-GB allocations), then the measured pass on warm buffers, as the
    # plain path after it runs
    cold = batch.run_batched(s, d, offs, counts, store, budget_pairs=budget_pairs)
    res = batch.run_batched(s, d, offs, counts, store, budget_pairs=budget_pairs)
    if not (np.array_equal(cold.rows, res.rows) and np.array_equal(cold.digests, res.digests)):
        raise SystemExit("batched passes disagree")
    picks = list(range(0, counts.size, plain_stride))
    plain = batch.run_plain(s, d, offs, counts, store, instances=picks)
    same = bool(np.array_equal(plain.rows[picks], res.rows[picks]) and
                np.array_equal(plain.digests[picks], res.digests[picks]))
    scale = counts.size / len(picks)
    tuples = int(res.rows.sum())
    out = {"instances": int(counts.size), "graph_replay": graph, "edges": int(s.size), "result_tuples": tuples,
           "waves": len(res.waves), "seconds": round(res.seconds, 3), "seconds_cold": round(cold.seconds, 3),
           "tuples_per_s": round(tuples / res.seconds, 1),
           "plain_seconds_measured": round(plain.seconds, 3), "plain_instances_measured": len(picks),
           "plain_seconds_estimated_all": round(plain.seconds * scale, 3),
           "plain_tuples_per_s": round(tuples / (plain.seconds * scale), 1),
           "speedup_vs_plain": round(plain.seconds * scale / res.seconds, 3),
           "results_equal_plain": same, "max_iterations": int(res.iterations.max())}
    if full_log:
        out["wave_log"] = res.waves
        out["cold_wave_log"] = cold.waves
    store.close()
    del store
    torch.cuda.empty_cache()
    return out


def run_batched_mode(args) -> None:
    r = batched_summary(args.budget_pairs, args.plain_stride, full_log=True, graph=_graph_arg(args),
                        tuning=_tuning_arg(args))
    r["mode"] = "batched"
    print(json.dumps(r), flush=True)


def _class_latency(results, requests, bounds: list[float]) -> list[dict]:
    """p50 / p99 latency per size class (edge-count bounds) of a stream."""
    from sparkalog_amd import serving

    edges = {r.rid: r.src.size for r in requests}
    out = []
    for i in range(len(bounds) + 1):
        lo = bounds[i - 1] if i else 0
        hi = bounds[i] if i < len(bounds) else float("inf")
        lat = [1000.0 * (x.done - x.arrival) for x in results if lo <= edges[x.rid] < hi]
        out.append({"edges_below": None if hi == float("inf") else hi, "requests": len(lat),
                    "p50_ms": round(serving.percentile(lat, 50), 3) if lat else None,
                    "p99_ms": round(serving.percentile(lat, 99), 3) if lat else None})
    return out


def served_summary(budget_pairs: float, requests: int, utilizations: str, lane_bounds: str, lane_slots: str,
                   one_lane: bool = True, graph=None, tuning="file", route: str = "edges") -> dict:
    """Served fixpoints: a Poisson stream of TC requests (edges log-uniform
    on [2^8, 2^20]) at arrival rates giving the plain path each requested
    utilisation (its mean service time measured on the same requests served
    back to back); continuous batching in size-class lanes (and, optionally,
    in one lane) against sequential plain execution of the same stream."""
    import torch

    from sparkalog_amd import serving
    from sparkalog_amd.execution import HipStream
    from sparkalog_amd.recursion import RelationStore

    torch.cuda.set_device(0)
    store = RelationStore(HipStream(), tuning=tuning, profile="batched")
    if graph is not None:
        store.set_graph(graph)
    graph = bool(store.tuning.get("graph", 0)) if graph is None else graph
    # the plain path's service time on these very requests (same seed: the
    # sizes do not depend on the rate), all present at the start
    back_to_back = serving.make_requests(requests, rate_per_s=1e9)
    serving.serve_plain(back_to_back[:4], store)  # warm-up
    _, secs = serving.serve_plain(back_to_back, store)
    mean_service = secs / requests
    bounds = [float(b) for b in lane_bounds.split(",")] if lane_bounds else []
    lanes = [RelationStore(HipStream(torch.cuda.Stream()), tuning=tuning, profile="batched")
             for _ in range(len(bounds) + 1)]
    for lane in lanes:
        lane.set_graph(graph)
    slots = [int(x) for x in lane_slots.split(",")] if lane_slots else None
    # warm-up: the whole stream once, so the lanes' stores have grown their
    # workspaces before anything is timed (as the plain store has)
    serving.serve_lanes(back_to_back, lanes, bounds, budget_pairs=budget_pairs, lane_slots=slots, route=route)
    out = {"requests": requests, "graph_replay": graph, "mean_plain_service_ms": round(1000 * mean_service, 3),
           "lane_route": route, "lane_bounds": bounds, "lane_slots": slots, "loads": []}
    equal = True
    for u in [float(x) for x in utilizations.split(",")]:
        rate = u / mean_service
        reqs = serving.make_requests(requests, rate_per_s=rate)
        mixed, t_mixed = serving.serve_lanes(reqs, lanes, bounds, budget_pairs=budget_pairs, lane_slots=slots,
                                             route=route)
        alone, t_alone = serving.serve_plain(reqs, store)
        equal = equal and all((m.rows, m.digest) == (a.rows, a.digest) for m, a in zip(mixed, alone))
        load = {"utilization": u, "rate_per_s": round(rate, 3), "lanes": serving.summarize(mixed, t_mixed),
                "plain": serving.summarize(alone, t_alone)}
        if route == "edges":
            load["lanes_by_class"] = _class_latency(mixed, reqs, bounds)
            load["plain_by_class"] = _class_latency(alone, reqs, bounds)
        if one_lane:
            one, t_one = serving.serve_batched(reqs, store, budget_pairs=budget_pairs)
            equal = equal and all((m.rows, m.digest) == (a.rows, a.digest) for m, a in zip(one, alone))
            load["one_lane"] = serving.summarize(one, t_one)
        out["loads"].append(load)
    out["results_equal_alone_vs_mix"] = equal
    store.close()
    for lane in lanes:
        lane.close()
    del store, lanes
    torch.cuda.empty_cache()
    return out


def run_served_mode(args) -> None:
    r = served_summary(args.budget_pairs, args.requests, args.utilization, args.lane_bounds, args.lane_slots,
                       graph=_graph_arg(args), tuning=_tuning_arg(args), route=args.lane_route)
    r["mode"] = "served"
    print(json.dumps(r), flush=True)


def _median(xs):
    xs = sorted(xs)
    return xs[(len(xs) - 1) // 2]


def run_crossover(args) -> None:
    """Placement crossover (the methodology of the reference's
    crates/sparkalog-cli/src/bin/*-crossover.rs): for each operator and row
    count, the oracle's serial and parallel CPU operators (the reference's
    CPU code restated; index built outside the timing), the HIP operator
    through the C ABI with its input resident in HBM (wall time of the call
    and its synchronization, as the reference times `*_cuda` + wait), and
    the PCIe copies of the input in either direction (pinned host memory).
    Every GPU result is checked against the oracle first. Prints one JSON
    line per case and a final line with the policies derived for MI355X
    (sparkalog_amd.placement.derive)."""
    import ctypes

    import numpy as np
    import torch

    import oracle
    from sparkalog_amd import opbench, placement, workload

    torch.cuda.set_device(0)
    threads = min(_cpu_threads(), CPU_SHARE)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    table = {}
    for name in args.crossover_ops.split(","):
        rows_list = [1 << k for k in range(args.crossover_min_log2, args.crossover_max_log2 + 1, 2)]
        t = {k: [] for k in ("rows", "cpu_serial_ms", "cpu_parallel_ms", "gpu_ms", "h2d_ms", "d2h_ms")}
        for n in rows_list:
            host = opbench.host_inputs(name, n, workload.KEY_UNIFORM)
            op = opbench.make(name, n, workload.KEY_UNIFORM, host=host)
            if name == "join":
                op.sync_after_count = True  # join_cuda reads the total before the emit (lib.rs:1608)
                index = oracle.range_index(host["right"][0])

            def cpu(th):
                if name == "filter":
                    return oracle.filter_rows(host["values"], op.OP_GE, op.OPERAND, threads=th)
                if name == "join":
                    return oracle.join(host["left"], host["right"], 1, 0, ((0, 0), (1, 1)), threads=th, index=index)
                return oracle_operator(name, host, th)

            op.launch(stream)
            torch.cuda.synchronize()
            got = op.result()
            exp = cpu(threads)
            ok = np.array_equal(got, exp) if name == "filter" else \
                (np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1]))
            if not ok:
                print(json.dumps({"op": name, "rows": n, "error": "result differs"}), flush=True)
                raise SystemExit(1)
            samples = args.samples if n <= (1 << 20) else 3
            ser, par, gpu, h2d, d2h = [], [], [], [], []
            for _ in range(samples):
                t0 = time.perf_counter()
                cpu(1)
                ser.append(1000 * (time.perf_counter() - t0))
                t0 = time.perf_counter()
                cpu(threads)
                par.append(1000 * (time.perf_counter() - t0))
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                op.launch(stream)
                torch.cuda.synchronize()
                gpu.append(1000 * (time.perf_counter() - t0))
            # the input's bytes across PCIe, from and to pinned host memory
            nbytes = int(opbench.input_bytes(name, host))
            pinned = torch.empty(max(nbytes // 4, 1), dtype=torch.int32, pin_memory=True)
            devbuf = torch.empty(max(nbytes // 4, 1), dtype=torch.int32, device="cuda")
            for _ in range(samples):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                devbuf.copy_(pinned, non_blocking=True)
                torch.cuda.synchronize()
                h2d.append(1000 * (time.perf_counter() - t0))
                t0 = time.perf_counter()
                pinned.copy_(devbuf, non_blocking=True)
                torch.cuda.synchronize()
                d2h.append(1000 * (time.perf_counter() - t0))
            measure = n + (op.right_rows if name == "union" else 0)
            row = {"op": name, "rows": n, "policy_rows": measure, "checked": True,
                   "cpu_serial_ms": round(_median(ser), 4), "cpu_parallel_ms": round(_median(par), 4),
                   "cpu_threads": threads, "gpu_ms": round(_median(gpu), 4), "h2d_ms": round(_median(h2d), 4),
                   "d2h_ms": round(_median(d2h), 4), "input_bytes": nbytes}
            print(json.dumps(row), flush=True)
            t["rows"].append(measure)
            for k in ("cpu_serial_ms", "cpu_parallel_ms", "gpu_ms", "h2d_ms", "d2h_ms"):
                t[k].append(row[k])
            del op, host, got, exp, pinned, devbuf
            torch.cuda.empty_cache()
        table[name] = t
    derived = placement.derive(table)
    print(json.dumps({"mode": "crossover", "cpu_threads": threads, "cpu_model": _cpu_model(),
                      "nproc": os.cpu_count(),
                      "derived_mi355x": {k: vars(v) for k, v in derived.items()},
                      "gb10": {k: vars(v) for k, v in placement.MEASURED_GB10.items()}}), flush=True)


def run_graph_mode(args) -> None:
    """Per-iteration HIP graph replay against the stream executor on the
    closure of a chain of --chain nodes (chain - 1 iterations of one new row
    per source: launch- and synchronization-bound), and on the 2^14-node
    component graph; results checked equal."""
    import numpy as np
    import torch

    from sparkalog_amd import plans, workload
    from sparkalog_amd.execution import HipStream
    from sparkalog_amd.recursion import RecursiveExecutor, RelationStore

    torch.cuda.set_device(0)
    R = plans.RelationId
    plan = plans.transitive_closure_scc(R(1), R(0))
    n = args.chain
    chain = (np.arange(n - 1, dtype=np.uint32), np.arange(1, n, dtype=np.uint32))
    s, d, _ = workload.component_graph(nodes=1 << 14)
    out = {"mode": "graph", "cases": []}
    for name, edges in ((f"chain_{n}", chain), ("component_graph_16384", (s, d))):
        case = {"workload": name}
        for graph in (0, 1):
            store = RelationStore(HipStream(), tuning=None)
            store.set_graph(bool(graph))

            def once():
                store.clear()
                store.insert_static(R(0), edges)
                store.insert_recursive(R(1), edges, edges)
                summ = RecursiveExecutor.compile(plan, store).run(store, 1_000_000)
                return summ, store.digest(R(1))

            once()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = max(1, args.steps)
            for _ in range(reps):
                summ, dig = once()
            torch.cuda.synchronize()
            ms = 1000.0 * (time.perf_counter() - t0) / reps
            caps, reps_g = store.graph_stats()
            upd, cap_ms = store.graph_capture_stats()
            case["graph" if graph else "stream"] = {
                "ms": round(ms, 3), "iterations": summ.iterations, "us_per_iteration": round(1000 * ms / summ.iterations, 2),
                "rows": dig[2], "digest": dig[0], "captures": caps, "replays": reps_g, "updates": upd,
                "capture_ms": round(cap_ms, 2)}
            store.close()
        case["results_equal"] = (case["graph"]["rows"], case["graph"]["digest"]) == \
            (case["stream"]["rows"], case["stream"]["digest"])
        out["cases"].append(case)
    print(json.dumps(out), flush=True)


def run_tune_mode(args) -> None:
    """Autotuner (sparkalog_amd.autotune): coordinate descent over the
    executor's knobs on each profile's calibration draw; writes the tuning
    file and prints one JSON line per profile (the sweep's trace stays in
    the file)."""
    import torch

    from sparkalog_amd import autotune

    torch.cuda.set_device(0)
    profiles = tuple(p for p in args.tune_profiles.split(",") if p)
    result = autotune.tune(profiles, log=lambda m: print(m, flush=True))
    out = Path(args.tuning_out)
    out.parent.mkdir(parents=True, exist_ok=True)
    autotune.write(result, out)
    for name, prof in result["profiles"].items():
        print(json.dumps({"mode": "tune", "profile": name, "file": str(out),
                          **{k: v for k, v in prof.items() if k != "trace"}}), flush=True)


def run_copy_mode(args) -> None:
    """The library's plain device copy of --copy-bytes, launched --steps
    times and timed with HIP events on its stream: the measured HBM copy
    ceiling, and (under rocprofv3 --pmc) the known-size kernel whose
    FETCH_SIZE / WRITE_SIZE readings scale those of the fixpoint's kernels."""
    import torch

    from sparkalog_amd.execution import HipStream, copy_u64, fill_mod_u32

    torch.cuda.set_device(0)
    stream = HipStream()
    words = args.copy_bytes // 8
    src = torch.empty(words * 2, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    fill_mod_u32(src, 1_000_003, stream)
    copy_u64(src, dst, stream)  # warm-up
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record(stream.torch_stream)
    for _ in range(args.steps):
        copy_u64(src, dst, stream)
    stop.record(stream.torch_stream)
    stop.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    ok = bool(torch.equal(dst, src))
    print(json.dumps({"mode": "copy", "kernel": "copy_u64_kernel", "bytes": words * 8, "launches": args.steps + 1,
                      "ms_per_copy": round(ms, 4), "gbs_read_plus_write": round(2 * words * 8 / ms / 1e6, 1),
                      "frac_of_hbm_peak": round(2 * words * 8 / ms / 1e6 / HBM_PEAK_GBS, 4), "equal": ok}),
          flush=True)


# The box's CPU share for one GPU: worker pools are sized to it whatever
# nproc reports there.
CPU_SHARE = 16


def _oracle_tc(nodes: int, threads: int) -> tuple[float, int, int, int]:
    """One oracle semi-naive fixpoint of the component graph of `nodes`
    nodes: (seconds, edges, result tuples, iterations)."""
    import oracle
    from sparkalog_amd import plans, workload

    s, d, _ = workload.component_graph(nodes=nodes)
    store = oracle.Store()
    store.insert_static(0, (s, d))
    store.insert_recursive(1, (s, d), (s, d))
    R = plans.RelationId
    rule = oracle.rule_from_plan(plans.transitive_closure_scc(R(1), R(0)).rules[0])
    t0 = time.perf_counter()
    rc, summary = store.run([1], [rule], 10_000, threads)
    dt = time.perf_counter() - t0
    if rc != 0:
        raise RuntimeError(f"oracle run failed with {rc}")
    return dt, int(s.size), int(store.get(1)[0].size), int(summary.iterations)


def cpu_baseline(sample_nodes: int, serial_nodes: int) -> dict:
    """The oracle's semi-naive executor (the reference's CPU operators and
    RecursiveExecutor restated in C) on this host: parallel (Rayon-style
    chunks, one worker per core of the GPU's CPU share) on a draw of
    `sample_nodes` nodes of the same component-graph distribution, and serial
    on `serial_nodes` nodes. The full 2^20-node graph (1.03e9 tuples) would
    take minutes on either, so bounded samples stand in for it."""
    nproc = os.cpu_count() or 1
    threads = min(_cpu_threads(), CPU_SHARE)
    dt, edges, rows, iters = _oracle_tc(sample_nodes, threads)
    sdt, sedges, srows, siters = _oracle_tc(serial_nodes, 1)
    return {"value": round(rows / dt, 1), "unit": "tuples/s", "cores": threads, "kind": "port",
            "sample": f"parallel: component graph of {sample_nodes} nodes (same generator, seed 0xC2; {edges} "
                      f"edges, {rows} result tuples, {iters} iterations) in {dt:.2f} s on {threads} threads",
            "serial": {"value": round(srows / sdt, 1), "unit": "tuples/s", "cores": 1,
                       "sample": f"component graph of {serial_nodes} nodes ({sedges} edges, {srows} result "
                                 f"tuples, {siters} iterations) in {sdt:.2f} s"},
            "host": {"nproc": nproc, "affinity": _cpu_threads(), "cpu_share_used": threads, "model": _cpu_model()}}


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "components and instances are independent, so each GPU would run its own share", file=sys.stderr)
        sys.exit(2)
    if args.mode == "ops":
        run_ops(args)
        return
    if args.mode == "batched":
        run_batched_mode(args)
        return
    if args.mode == "served":
        run_served_mode(args)
        return
    if args.mode == "copy":
        run_copy_mode(args)
        return
    if args.mode == "crossover":
        run_crossover(args)
        return
    if args.mode == "tune":
        run_tune_mode(args)
        return
    if args.mode == "graph":
        run_graph_mode(args)
        return
    import numpy as np
    import torch

    from sparkalog_amd import autotune, plans, workload
    from sparkalog_amd.execution import HipStream
    from sparkalog_amd.recursion import RecursiveExecutor, RelationStore

    torch.cuda.set_device(0)
    src, dst, comps = workload.component_graph(nodes=args.nodes)
    # inputs resident in HBM before timing
    d_src = torch.from_numpy(src.view(np.int32)).cuda()
    d_dst = torch.from_numpy(dst.view(np.int32)).cuda()
    stream = HipStream()
    store = RelationStore(stream, tuning=_tuning_arg(args))
    overrides = {k: v for k, v in (("merge_ratio", args.merge_ratio), ("indexed_anti_join", args.indexed_anti_join),
                                   ("sparse_probe_rows", args.sparse_probe_rows)) if v is not None}
    autotune.apply(store, overrides)
    tuned_values = {**store.tuning, **overrides}
    R = plans.RelationId
    plan = plans.transitive_closure_scc(R(1), R(0))

    def step(profile: bool, on=None):
        st = store if on is None else on
        st.clear()
        st.insert_static(R(0), (d_src, d_dst))
        st.insert_recursive(R(1), (d_src, d_dst), (d_src, d_dst))
        ex = RecursiveExecutor.compile(plan, st)
        return ex.run(st, 10_000)

    for _ in range(args.warmup):
        step(False)
    # the timed steps run without per-operator events (recording them adds
    # about 3% to a fixpoint); as many profiled steps follow, timed too, for
    # the per-operator breakdown and the roofline
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    summaries = [step(False) for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    store.set_profiling(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    profiled = [step(True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed_profiled = time.perf_counter() - t0
    ops = store.op_stats()
    store.set_profiling(False)
    if profiled[-1].relation_rows != summaries[-1].relation_rows:
        raise SystemExit("profiled steps disagree with the timed steps")

    rows = summaries[-1].relation_rows[0][1]
    ms_per_step = 1000.0 * elapsed / args.steps
    tuples_per_s = rows / (ms_per_step / 1000.0)
    algo_bytes = summaries[-1].algorithmic_bytes
    device_ms = sum(s.device_ms for s in summaries) / len(summaries)

    # dominant kernel: the operator main kernel with the most device time
    # over the profiled steps (HIP events around that kernel alone); its
    # algorithmic bytes are the operator's (8 bytes per pair read or written)
    timed = [(k, v) for k, v in ops.items() if v["kernel_launches"] and v["kernel_launches"] == v["launches"]]
    dom_op, dom = max(timed, key=lambda kv: kv[1]["kernel_ms"])
    launches = dom["kernel_launches"]
    avg_ms = dom["kernel_ms"] / launches
    bytes_per_launch = dom["bytes"] / launches
    achieved = bytes_per_launch / (avg_ms / 1000.0) / 1e9
    traffic = None
    if PROFILE_SUMMARY.exists():
        prof = json.loads(PROFILE_SUMMARY.read_text())
        # a kernel chain ("a+b+c", timed as one) moves the sum of its kernels' bytes
        per_kernel = prof.get("traffic_bytes_per_launch", {})
        parts = dom["kernel"].split("+")
        if all(k in per_kernel for k in parts):
            traffic = sum(per_kernel[k] for k in parts)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom["kernel"],
                "operator": dom_op, "launches_per_step": launches // args.steps,
                "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": round(bytes_per_launch, 1)}

    line = {
        "metric": "TC result tuples/sec + HBM GB/s (1 GPU); batched/served p50/p99 vs plain",
        "value": round(tuples_per_s, 1),
        "unit": "tuples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "ms_per_step_profiled": round(1000.0 * elapsed_profiled / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "single-instance transitive closure, component graph",
                   "nodes": args.nodes, "edges": int(src.size), "components": int(comps),
                   "result_tuples": int(rows), "iterations": int(summaries[-1].iterations)},
        "hbm_gbs": round(algo_bytes / (ms_per_step / 1000.0) / 1e9, 1),
        "device_ms_per_step": round(device_ms, 3),
        "ops": {k: {"launches": v["launches"] // args.steps, "ms": round(v["ms"] / args.steps, 3),
                    "kernel_ms": round(v["kernel_ms"] / args.steps, 3),
                    "gbs": round(v["bytes"] / (v["ms"] / 1000.0) / 1e9, 1) if v["ms"] else None}
                for k, v in ops.items() if v["launches"]},
        "roofline": roofline,
    }
    tuning_file = autotune.default_file() if args.tuning == "file" else (None if args.tuning == "none"
                                                                          else Path(args.tuning))
    line["tuning"] = {"file": tuning_file.name if tuning_file else None, "profile": "closure",
                      "values": tuned_values}
    if tuned_values and not args.no_default_compare:
        # the same fixpoint under the built-in defaults, on a store of its own
        base = RelationStore(stream, tuning=None)
        step(False, base)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        base_summ = [step(False, base) for _ in range(max(2, args.steps))]
        torch.cuda.synchronize()
        base_ms = 1000.0 * (time.perf_counter() - t0) / len(base_summ)
        line["tuning"].update({
            "default_ms_per_step": round(base_ms, 3), "tuned_ms_per_step": round(ms_per_step, 3),
            "speedup_vs_default": round(base_ms / ms_per_step, 4),
            "results_equal_default": base_summ[-1].relation_rows == summaries[-1].relation_rows and
            base.digest(R(1)) == store.digest(R(1))})
        base.close()
    if args.host_io:
        # edges uploaded from host arrays (pinned staging), fixpoint, closure
        # read back to host memory: the rate a caller handing over host
        # buffers sees
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        store.clear()
        store.insert_static(R(0), (src, dst))
        store.insert_recursive(R(1), (src, dst), (src, dst))
        RecursiveExecutor.compile(plan, store).run(store, 10_000)
        t1 = time.perf_counter()
        host_full = store.full(R(1))
        t2 = time.perf_counter()
        line["host_io"] = {"seconds": round(t2 - t0, 4), "upload_and_fixpoint_s": round(t1 - t0, 4),
                           "readback_s": round(t2 - t1, 4), "rows": int(host_full[0].size),
                           "tuples_per_s_pcie_inclusive": round(host_full[0].size / (t2 - t0), 1)}
        del host_full
    if not args.no_batched:
        # the batched configuration (4,096 instances) beside the plain path;
        # the closure's store returns its pools first (close, not only del:
        # other references may keep the object and its device memory alive)
        store.close()
        del store
        torch.cuda.empty_cache()
        line["batched"] = batched_summary(args.budget_pairs, args.bench_plain_stride, graph=_graph_arg(args),
                                          tuning=_tuning_arg(args))
    if not args.no_served:
        line["served"] = served_summary(args.budget_pairs, args.requests, args.bench_utilization, args.lane_bounds,
                                        args.lane_slots, one_lane=False, graph=_graph_arg(args), route=args.lane_route,
                                        tuning=_tuning_arg(args))
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_nodes, args.cpu_serial_nodes)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
