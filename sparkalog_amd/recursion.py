"""Device-resident semi-naive fixpoint (the HIP counterpart of
crates/sparkalog-recursion/src/generic.rs).

RelationStore keeps FULL/DELTA/NEWT of every relation in HBM;
RecursiveExecutor.compile validates a RecursiveSccPlan exactly as the
reference does (generic.rs:407-470) and run() iterates on the device until
every DELTA is empty (generic.rs:472-516).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from .plans import RecursiveSccPlan, RelationId

# sparkalog_hip_status codes
OK = 0
STATUS_NAMES = {
    1: "Hip", 2: "TooManyRows", 3: "MissingPlanColumn", 4: "IterationLimit",
    5: "IncompatibleTargetPlans", 6: "NotRecursive", 7: "Missing", 8: "Duplicate",
    9: "EmptyScc", 10: "DeltaOutsideScc", 11: "TargetOutsideScc", 12: "MissingRuleTarget",
    13: "OutOfMemory", 14: "UnsupportedPlan", 15: "DeviceError", 16: "InvalidArgument",
}


class GenericExecutionError(Exception):
    """GenericExecutionError / RelationStoreError (generic.rs:129-135,
    :544-559); `kind` names the variant."""

    def __init__(self, status: int, detail: str = ""):
        self.status = status
        self.kind = STATUS_NAMES.get(status, f"Status{status}")
        text = _native.hip().sparkalog_hip_status_string(status).decode()
        if status == 1:
            text += f" (hipError {_native.hip().sparkalog_hip_last_hip_error()})"
        super().__init__(f"{self.kind}: {text}{detail}")


def _check(status: int, detail: str = "") -> None:
    if status != OK:
        raise GenericExecutionError(status, detail)


@dataclass
class SccSummary:
    """SccSummary (generic.rs:536-542) plus device time and traffic."""

    iterations: int
    total_new_rows: int
    relation_rows: list[tuple[RelationId, int]]
    device_ms: float
    algorithmic_bytes: float


def _host_u32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _columns(rel):
    """(c0, c1, rows, on_device, keepalive) from numpy-likes or device tensors."""
    a, b = rel
    if hasattr(a, "is_cuda") and a.is_cuda:
        return (ctypes.c_void_p(a.data_ptr()) if a.numel() else None,
                ctypes.c_void_p(b.data_ptr()) if b.numel() else None, int(a.numel()), 1, (a, b))
    a, b = _host_u32(a), _host_u32(b)
    if a.shape != b.shape:
        raise ValueError("columns differ in length")
    return (a.ctypes.data_as(ctypes.c_void_p) if a.size else None,
            b.ctypes.data_as(ctypes.c_void_p) if b.size else None, int(a.size), 0, (a, b))


class RelationStore:
    """RelationStore (generic.rs:16-127) in device memory."""

    def __init__(self, stream=None, tuning="file", profile: str = "closure"):
        """`tuning`: "file" = the autotuner's file (autotune.default_file(),
        if any), None = built-in defaults, or a dict / tuning file path;
        `profile` names the file's profile ("closure" or "batched"). Results
        never depend on it."""
        from . import autotune

        self._lib = _native.hip()
        handle = ctypes.c_void_p()
        raw = None if stream is None else getattr(stream, "raw", stream)
        _check(self._lib.sparkalog_hip_store_create(ctypes.byref(handle), raw))
        self._h = handle
        if isinstance(tuning, str) and tuning == "file":
            tuning = autotune.default_file()
        self.tuning = autotune.profile_values(tuning, profile)
        autotune.apply(self, self.tuning)

    def close(self) -> None:
        if self._h:
            self._lib.sparkalog_hip_store_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def insert_static(self, id: RelationId, relation) -> None:
        c0, c1, n, dev, _keep = _columns(relation)
        _check(self._lib.sparkalog_hip_store_insert_static(self._h, int(id), c0, c1, n, dev))

    def insert_recursive(self, id: RelationId, full, delta) -> None:
        f0, f1, nf, dev_f, _k1 = _columns(full)
        d0, d1, nd, dev_d, _k2 = _columns(delta)
        if dev_f != dev_d:
            raise ValueError("FULL and DELTA must both be host or both be device columns")
        _check(self._lib.sparkalog_hip_store_insert_recursive(self._h, int(id), f0, f1, nf, d0, d1, nd, dev_f))

    def clear(self) -> None:
        """Forget every relation; device buffers stay allocated for reuse."""
        _check(self._lib.sparkalog_hip_store_clear(self._h))

    def rows(self, id: RelationId, which: int = 0) -> int:
        out = ctypes.c_uint64(0)
        _check(self._lib.sparkalog_hip_store_rows(self._h, int(id), which, ctypes.byref(out)))
        return out.value

    def _read(self, id: RelationId, which: int) -> tuple[np.ndarray, np.ndarray]:
        n = self.rows(id, which)
        a = np.empty(max(n, 1), dtype=np.uint32)
        b = np.empty(max(n, 1), dtype=np.uint32)
        _check(self._lib.sparkalog_hip_store_read(self._h, int(id), which, a.ctypes.data_as(ctypes.c_void_p),
                                                  b.ctypes.data_as(ctypes.c_void_p), n, 0))
        return a[:n], b[:n]

    def full(self, id: RelationId) -> tuple[np.ndarray, np.ndarray]:
        return self._read(id, 0)

    def full_device(self, id: RelationId, device=None):
        """FULL copied into two new device tensors (int32 holding u32)."""
        import torch

        n = self.rows(id, 0)
        dev = device or torch.device("cuda", torch.cuda.current_device())
        a = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        b = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        _check(self._lib.sparkalog_hip_store_read(self._h, int(id), 0, ctypes.c_void_p(a.data_ptr()),
                                                  ctypes.c_void_p(b.data_ptr()), n, 1))
        return a[:n], b[:n]

    def delta(self, id: RelationId) -> tuple[np.ndarray, np.ndarray]:
        return self._read(id, 1)

    def digest(self, id: RelationId) -> tuple[int, bool, int]:
        """(order-independent digest, strictly sorted?, rows) of FULL."""
        d = ctypes.c_uint64(0)
        ok = ctypes.c_int(0)
        n = ctypes.c_uint64(0)
        _check(self._lib.sparkalog_hip_store_digest(self._h, int(id), ctypes.byref(d), ctypes.byref(ok), ctypes.byref(n)))
        return d.value, bool(ok.value), n.value

    # ---- many instances in one relation (batched and served fixpoints) ----

    def set_instance_tracking(self, instance_bits: int, instances: int) -> None:
        _check(self._lib.sparkalog_hip_set_instance_tracking(self._h, instance_bits, instances))

    def instance_iterations(self, instances: int) -> np.ndarray:
        out = np.zeros(max(instances, 1), dtype=np.uint32)
        _check(self._lib.sparkalog_hip_instance_iterations(self._h, out.ctypes.data_as(ctypes.c_void_p), instances))
        return out[:instances]

    def instance_delta_rows(self, instances: int) -> np.ndarray:
        """Per instance, DELTA rows after the last step (read back with the
        step's counters: no device work here)."""
        out = np.zeros(instances, np.uint64)
        _check(self._lib.sparkalog_hip_instance_delta_rows(self._h, out.ctypes.data_as(ctypes.c_void_p), instances))
        return out

    def instance_stats(self, id: RelationId, instance_bits: int, instances: int) -> tuple[np.ndarray, np.ndarray]:
        rows = np.zeros(max(instances, 1), dtype=np.uint64)
        digests = np.zeros(max(instances, 1), dtype=np.uint64)
        _check(self._lib.sparkalog_hip_store_instance_stats(self._h, int(id), instance_bits, instances,
                                                            rows.ctypes.data_as(ctypes.c_void_p),
                                                            digests.ctypes.data_as(ctypes.c_void_p)))
        return rows[:instances], digests[:instances]

    def range_stats(self, id: RelationId, which: int, first_lo: int, first_hi: int, base: int) -> tuple[int, int]:
        rows = ctypes.c_uint64(0)
        digest = ctypes.c_uint64(0)
        _check(self._lib.sparkalog_hip_store_range_stats(self._h, int(id), which, first_lo, first_hi, base,
                                                         ctypes.byref(rows), ctypes.byref(digest)))
        return rows.value, digest.value

    def range_rows(self, id: RelationId, which: int, bounds) -> np.ndarray:
        """Rows per first-column range [bounds[2j], bounds[2j+1])."""
        b = np.ascontiguousarray(np.asarray(bounds, dtype=np.uint32))
        n = b.size // 2
        out = np.zeros(max(n, 1), dtype=np.uint64)
        _check(self._lib.sparkalog_hip_store_range_rows(self._h, int(id), which, b.ctypes.data_as(ctypes.c_void_p), n,
                                                        out.ctypes.data_as(ctypes.c_void_p)))
        return out[:n]

    def merge_rows(self, id: RelationId, which: int, relation) -> None:
        c0, c1, n, dev, _keep = _columns(relation)
        _check(self._lib.sparkalog_hip_store_merge_rows(self._h, int(id), which, c0, c1, n, dev))

    def difference(self, target: RelationId, left: RelationId, right: RelationId) -> int:
        """target := FULL(left) \\ FULL(right) as a static relation (a
        non-recursive rule with stratified negation); returns its rows."""
        n = ctypes.c_uint64(0)
        _check(self._lib.sparkalog_hip_store_difference(self._h, int(target), int(left), int(right), ctypes.byref(n)))
        return n.value

    def remove_range(self, id: RelationId, first_lo: int, first_hi: int) -> None:
        _check(self._lib.sparkalog_hip_store_remove_range(self._h, int(id), first_lo, first_hi))

    def remove_ranges(self, id: RelationId, bounds) -> None:
        """Remove every row whose first column is in one of the [lo, hi)
        ranges (a flat array lo0, hi0, lo1, hi1, ... ascending, disjoint)."""
        b = np.ascontiguousarray(bounds, dtype=np.uint32)
        if b.size % 2:
            raise ValueError("remove_ranges: bounds come in (lo, hi) pairs")
        _check(self._lib.sparkalog_hip_store_remove_ranges(self._h, int(id), b.ctypes.data_as(_native.u32p),
                                                           b.size // 2))

    def set_tuning(self, merge_ratio: int = 3, segmented_distinct: bool = True, indexed_anti_join: bool = False,
                   sparse_probe_rows: int = 1 << 22) -> None:
        """Run-merge ratio of FULL's sorted runs, the grouped (per-run LDS)
        distinct, the first-column-indexed anti-join and the candidate count
        up to which FULL is probed one candidate per thread; results are
        identical for every setting."""
        _check(self._lib.sparkalog_hip_store_set_tuning(self._h, merge_ratio, int(segmented_distinct),
                                                        int(indexed_anti_join), sparse_probe_rows))

    JOIN_INDEX = {"auto": 0, "dense": 1, "hashed": 2}

    def set_option(self, name: str, value: int) -> None:
        """A named executor option (sparkalog_hip_store_set_option); results
        do not depend on any of them."""
        _check(self._lib.sparkalog_hip_store_set_option(self._h, name.encode(), int(value)), name)

    def set_join_index(self, layout: str) -> None:
        """Layout of the join index: "auto", "dense" or "hashed"."""
        self.set_option("join_index", self.JOIN_INDEX[layout])

    def set_graph(self, enabled: bool) -> None:
        """Replay each iteration of single-rule fixpoints from a captured HIP
        graph (one host synchronization per iteration)."""
        self.set_option("graph", int(bool(enabled)))

    def graph_stats(self) -> tuple[int, int]:
        """(graphs captured, graph replays) so far."""
        c = ctypes.c_uint64(0)
        r = ctypes.c_uint64(0)
        _check(self._lib.sparkalog_hip_store_graph_stats(self._h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def graph_capture_stats(self) -> tuple[int, float]:
        """(captures that updated an executable graph in place, host ms spent
        capturing and instantiating or updating)."""
        u = ctypes.c_uint64(0)
        ms = ctypes.c_double(0)
        _check(self._lib.sparkalog_hip_store_graph_capture_stats(self._h, ctypes.byref(u), ctypes.byref(ms)))
        return u.value, ms.value

    def set_profiling(self, enabled: bool) -> None:
        _check(self._lib.sparkalog_hip_store_set_profiling(self._h, int(enabled)))

    def op_stats(self) -> dict[str, dict]:
        """Per operator: launches, device ms (helper kernels included),
        algorithmic bytes, and the main kernel's name, launches and ms."""
        out = {}
        i = 0
        while True:
            name, kernel = ctypes.c_char_p(), ctypes.c_char_p()
            k_launches, k_ms = ctypes.c_uint64(0), ctypes.c_double(0)
            if self._lib.sparkalog_hip_store_op_kernel_stats(self._h, i, ctypes.byref(name), ctypes.byref(kernel),
                                                             ctypes.byref(k_launches), ctypes.byref(k_ms)) != OK:
                break
            launches, ms, nbytes = ctypes.c_uint64(0), ctypes.c_double(0), ctypes.c_double(0)
            _check(self._lib.sparkalog_hip_store_op_stats(self._h, i, ctypes.byref(launches), ctypes.byref(ms),
                                                          ctypes.byref(nbytes)))
            out[name.value.decode()] = {"launches": launches.value, "ms": ms.value, "bytes": nbytes.value,
                                        "kernel": kernel.value.decode() or None, "kernel_launches": k_launches.value,
                                        "kernel_ms": k_ms.value}
            i += 1
        return out


def _rule_struct(rule) -> _native.Rule:
    r = _native.Rule()
    r.target, r.delta_input, r.right_input = int(rule.target), int(rule.delta_input), int(rule.right_input)
    r.left_key, r.right_key = rule.join.left_key, rule.join.right_key
    for i in range(2):
        r.out_input[i] = int(rule.join.output[i].input)
        r.out_column[i] = rule.join.output[i].column
        r.distinct_columns[i] = rule.distinct.columns[i]
        r.anti_left[i], r.anti_right[i] = rule.anti_join.left[i], rule.anti_join.right[i]
        r.union_left[i], r.union_right[i] = rule.union.left[i], rule.union.right[i]
    return r


class RecursiveExecutor:
    """RecursiveExecutor (generic.rs:400-534) on the device."""

    def __init__(self, plan: RecursiveSccPlan):
        self.plan = plan
        self._relations = (ctypes.c_uint32 * max(1, len(plan.relations)))(*[int(r) for r in plan.relations])
        self._rules = (_native.Rule * max(1, len(plan.rules)))(*[_rule_struct(r) for r in plan.rules])

    @classmethod
    def compile(cls, plan: RecursiveSccPlan, store: RelationStore) -> "RecursiveExecutor":
        ex = cls(plan)
        _check(_native.hip().sparkalog_hip_executor_compile(store._h, ex._relations, len(plan.relations),
                                                            ex._rules, len(plan.rules)))
        return ex

    def begin(self, store: RelationStore) -> None:
        """Stepwise run: compile and start (see step/end)."""
        _check(_native.hip().sparkalog_hip_session_begin(store._h, self._relations, len(self.plan.relations),
                                                         self._rules, len(self.plan.rules)))

    def step(self, store: RelationStore) -> int:
        """One semi-naive iteration (none when every DELTA is empty); returns the
        DELTA rows left."""
        left = ctypes.c_uint64(0)
        _check(_native.hip().sparkalog_hip_session_step(store._h, ctypes.byref(left)))
        return left.value

    def end(self, store: RelationStore) -> SccSummary:
        summary = _native.Summary()
        _check(_native.hip().sparkalog_hip_session_end(store._h, ctypes.byref(summary)))
        return SccSummary(iterations=summary.iterations, total_new_rows=summary.total_new_rows,
                          relation_rows=[(r, store.rows(r, 0)) for r in self.plan.relations],
                          device_ms=summary.device_ms, algorithmic_bytes=summary.algorithmic_bytes)

    def run(self, store: RelationStore, max_iterations: int) -> SccSummary:
        summary = _native.Summary()
        status = _native.hip().sparkalog_hip_executor_run(store._h, self._relations, len(self.plan.relations),
                                                          self._rules, len(self.plan.rules), max_iterations,
                                                          ctypes.byref(summary))
        if status == 4:
            raise GenericExecutionError(
                status, f"; {summary.delta_rows_left} delta rows remain after {max_iterations} iterations")
        _check(status)
        return SccSummary(
            iterations=summary.iterations,
            total_new_rows=summary.total_new_rows,
            relation_rows=[(r, store.rows(r, 0)) for r in self.plan.relations],
            device_ms=summary.device_ms,
            algorithmic_bytes=summary.algorithmic_bytes,
        )
