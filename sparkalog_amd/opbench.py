"""Operator drop-in microbench (the crossover-harness methodology of the
reference's crates/sparkalog-cli/src/bin/*-crossover.rs, on MI355X).

Each operator runs through the C ABI on device-resident u32 pair columns of
2^k rows. Inputs are produced outside the timed interval; every launch is
bracketed by HIP events on the launching stream; p10/median/p90 use the
reference's index rule `(len-1)*num/den` (join-crossover.rs:119-122). A
caller checks `result()` against an independent CPU result before it keeps
any timing (distinct-crossover.rs:227-234).

Workloads (keys are the join/sort-major column, uniform on [0, 2^24) or
Zipf(1.1) ranks):
  distinct   rows with 30% duplicate pairs
  union      two sorted unique relations of `rows` pairs sharing 50%
  anti_join  the same two relations, left minus right
  join       probe = `rows` pairs (value, key); build = 2^24 pairs (key,
             value) with uniform keys (about one match per probe); output
             (probe value, build value), the transitive-closure plan
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native, workload

KEY_SPACE = 1 << 24
OPERATORS = ("distinct", "union", "anti_join", "join")
JOIN_BUILD_ROWS = 1 << 24


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def _temp(nbytes: int) -> torch.Tensor:
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda")


def percentile(sorted_values, num, den):
    return sorted_values[(len(sorted_values) - 1) * num // den]


def range_index(keys: np.ndarray):
    """U32RangeIndex (crates/sparkalog-storage/src/lib.rs:357-393) of a key
    column: sorted unique keys, starts[K+1], row ids ordered by (key, row)."""
    order = np.argsort(keys, kind="stable").astype(np.uint32)
    unique, first = np.unique(keys[order], return_index=True)
    return unique.astype(np.uint32), np.append(first, keys.size).astype(np.uint32), order


def host_inputs(name: str, rows: int, key_mode: int, seed: int = workload.C3_SEED) -> dict:
    """The seeded host arrays of one operator case."""
    if name == "distinct":
        return {"input": workload.distinct_input(seed, rows, 0.7, key_mode, KEY_SPACE)}
    if name in ("union", "anti_join"):
        left, right = workload.overlapping_sets(seed, rows, 0.5, key_mode, KEY_SPACE)
        return {"left": left, "right": right}
    if name == "join":
        pk, pv = workload.pairs(seed, rows, key_mode, KEY_SPACE)
        bk, bv = workload.pairs(seed + 1, JOIN_BUILD_ROWS, workload.KEY_UNIFORM, KEY_SPACE)
        return {"left": (pv, pk), "right": (bk, bv), "index": range_index(bk)}
    if name == "filter":
        return {"values": (np.arange(rows, dtype=np.uint64) % Filter.MODULUS).astype(np.uint32)}
    raise ValueError(name)


def input_bytes(name: str, host: dict) -> float:
    """Bytes of the operator's input columns (what a CPU-produced input must
    cross PCIe with)."""
    if name == "filter":
        return 4.0 * host["values"].size
    if name == "distinct":
        return 8.0 * host["input"][0].size
    if name == "join":
        return 8.0 * host["left"][0].size
    return 8.0 * (host["left"][0].size + host["right"][0].size)


class _Op:
    name = ""

    def __init__(self, host: dict):
        self.host = host

    def algorithmic_bytes(self) -> float:
        raise NotImplementedError

    def launch(self, stream) -> None:
        raise NotImplementedError

    def out_rows(self) -> int:
        raise NotImplementedError

    def result(self):
        u = self.out_rows()
        return (self.o0[:u].cpu().numpy().view(np.uint32), self.o1[:u].cpu().numpy().view(np.uint32))


def _count(t: torch.Tensor, width: int) -> int:
    mask = (1 << width) - 1
    v = int(t.item()) & mask
    if v == mask:
        raise RuntimeError("a device tile wait gave up (all-ones output count)")
    return v


class Distinct(_Op):
    name = "distinct"

    def __init__(self, host: dict):
        super().__init__(host)
        a, b = host["input"]
        rows = a.size
        self.rows = rows
        self.a, self.b = _dev(a), _dev(b)
        self.packed = torch.empty(rows, dtype=torch.int64, device="cuda")
        self.scratch = torch.empty(rows, dtype=torch.int64, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.o0 = torch.empty(rows, dtype=torch.int32, device="cuda")
        self.o1 = torch.empty(rows, dtype=torch.int32, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_distinct_u32_temporary_bytes(_p(self.packed), _p(self.scratch), _p(self.count),
                                                                 rows, ctypes.byref(tb), None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_distinct_u32(_p(self.a), _p(self.b), self.rows, _p(self.packed),
                                                      _p(self.scratch), _p(self.count), _p(self.o0), _p(self.o1),
                                                      _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc

    def out_rows(self):
        return _count(self.count, 64)

    def algorithmic_bytes(self):
        return 8.0 * self.rows + 8.0 * self.out_rows()


class _SetOp(_Op):
    def __init__(self, host: dict):
        super().__init__(host)
        (lf, ls), (rf, rs) = host["left"], host["right"]
        self.rows = lf.size
        self.right_rows = rf.size
        self.lf, self.ls, self.rf, self.rs = _dev(lf), _dev(ls), _dev(rf), _dev(rs)


class Union(_SetOp):
    name = "union"

    def __init__(self, host: dict):
        super().__init__(host)
        n = self.rows + self.right_rows
        self.o0 = torch.empty(n, dtype=torch.int32, device="cuda")
        self.o1 = torch.empty(n, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_union_u32_temporary_bytes(None, self.rows, None, self.right_rows, None, None,
                                                              _p(self.count), ctypes.byref(tb), None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_union_u32(_p(self.lf), _p(self.ls), self.rows, _p(self.rf), _p(self.rs),
                                                   self.right_rows, None, None, None, None, _p(self.count),
                                                   _p(self.o0), _p(self.o1), _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc

    def out_rows(self):
        return _count(self.count, 64)

    def algorithmic_bytes(self):
        return 8.0 * (self.rows + self.right_rows) + 8.0 * self.out_rows()


class AntiJoin(_SetOp):
    name = "anti_join"

    def __init__(self, host: dict):
        super().__init__(host)
        self.o0 = torch.empty(self.rows, dtype=torch.int32, device="cuda")
        self.o1 = torch.empty(self.rows, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int32, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_anti_join_u32_temporary_bytes(None, None, _p(self.count), self.rows,
                                                                  ctypes.byref(tb), None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_anti_join_u32(_p(self.lf), _p(self.ls), self.rows, _p(self.rf), _p(self.rs),
                                                       self.right_rows, None, None, _p(self.count), _p(self.o0),
                                                       _p(self.o1), _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc

    def out_rows(self):
        return _count(self.count, 32)

    def algorithmic_bytes(self):
        return 8.0 * (self.rows + self.right_rows) + 8.0 * self.out_rows()


class Join(_Op):
    """probe (value, key) ⋈ build (key, value) on probe.1 = build.0 with
    output (probe.0, build.1) — the TC join plan."""

    name = "join"

    def __init__(self, host: dict):
        super().__init__(host)
        (pv, pk), (bk, bv) = host["left"], host["right"]
        keys, starts, order = host["index"]
        self.rows = pk.size
        self.build_rows = bk.size
        self.l0, self.lk = _dev(pv), _dev(pk)
        self.r0, self.r1 = _dev(bk), _dev(bv)
        self.keys, self.starts, self.irows = _dev(keys), _dev(starts), _dev(order)
        self.counts = torch.empty(self.rows, dtype=torch.int64, device="cuda")
        self.offsets = torch.empty(self.rows, dtype=torch.int64, device="cuda")
        self.total = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_join_u32_temporary_bytes(_p(self.counts), _p(self.offsets), self.rows,
                                                             ctypes.byref(tb), None)
        self.temp = _temp(tb.value)
        self.o0 = self.o1 = None
        self.sync_after_count = False

    def launch(self, stream):
        lib = _native.hip()
        rc = lib.sparkalog_hip_join_u32_count(_p(self.lk), self.rows, _p(self.keys), _p(self.starts),
                                              self.keys.numel(), _p(self.counts), _p(self.offsets), _p(self.total),
                                              _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc
        if self.o0 is None or self.sync_after_count:
            # the reference synchronizes here to size the output (lib.rs:1608)
            torch.cuda.synchronize()
        if self.o0 is None:
            m = self.out_rows()
            self.o0 = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
            self.o1 = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
        rc = lib.sparkalog_hip_join_u32_emit(_p(self.lk), self.rows, _p(self.keys), _p(self.starts), _p(self.irows),
                                             self.keys.numel(), _p(self.l0), 0, _p(self.r1), 1, _p(self.offsets),
                                             _p(self.o0), _p(self.o1), stream)
        assert rc == 0, rc

    def out_rows(self):
        return _count(self.total, 64)

    def algorithmic_bytes(self):
        return 8.0 * self.rows + 8.0 * self.build_rows + 8.0 * self.out_rows()


class Filter(_Op):
    """filter_cuda (crates/sparkalog-execution/src/lib.rs:1809-1875): ordered
    row ids of input >= operand; the crossover input is i % 97 with operand
    87 (about 10% of the rows, filter-crossover.rs's 10% case)."""

    name = "filter"
    MODULUS, OPERAND, OP_GE = 97, 87, 5

    def __init__(self, host: dict):
        super().__init__(host)
        v = host["values"]
        self.rows = v.size
        self.v = _dev(v)
        self.out = torch.empty(max(self.rows, 1), dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int32, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_filter_u32_temporary_bytes(None, _p(self.out), _p(self.count), self.rows,
                                                               ctypes.byref(tb), None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_filter_u32(_p(self.v), self.rows, self.OP_GE, self.OPERAND, None,
                                                    _p(self.out), _p(self.count), _p(self.temp), self.temp.numel(),
                                                    stream)
        assert rc == 0, rc

    def out_rows(self):
        return _count(self.count, 32)

    def algorithmic_bytes(self):
        return 4.0 * self.rows + 4.0 * self.out_rows()

    def result(self):
        return self.out[: self.out_rows()].cpu().numpy().view(np.uint32)


_CLASSES = {"distinct": Distinct, "union": Union, "anti_join": AntiJoin, "join": Join, "filter": Filter}


def make(name: str, rows: int, key_mode: int, seed: int = workload.C3_SEED, host: dict | None = None) -> _Op:
    """The device operator of one case (inputs uploaded, outside any timing)."""
    if name not in _CLASSES:
        raise ValueError(name)
    return _CLASSES[name](host if host is not None else host_inputs(name, rows, key_mode, seed))


def time_op(op: _Op, samples: int, warmup: int = 1) -> dict:
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for _ in range(warmup):
        op.launch(stream)
    torch.cuda.synchronize()
    times = []
    for _ in range(samples):
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        op.launch(stream)
        end.record()
        end.synchronize()
        times.append(start.elapsed_time(end))
    times.sort()
    med = percentile(times, 1, 2)
    nbytes = op.algorithmic_bytes()
    return {"op": op.name, "rows": op.rows, "out_rows": op.out_rows(), "samples": samples,
            "p10_ms": round(percentile(times, 1, 10), 4), "median_ms": round(med, 4),
            "p90_ms": round(percentile(times, 9, 10), 4), "algorithmic_bytes": nbytes,
            "gbs": round(nbytes / (med / 1000.0) / 1e9, 1)}
