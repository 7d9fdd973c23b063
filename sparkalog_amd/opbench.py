"""Operator drop-in microbench (the crossover-harness methodology of the
reference's crates/sparkalog-cli/src/bin/*-crossover.rs, on MI355X).

Each operator runs through the C ABI on device-resident u32 pair columns of
2^k rows. Inputs are produced outside the timed interval; every launch is
bracketed by HIP events on the launching stream; the first launch's output
is kept for checking; p10/median/p90 use the reference's index rule
`(len-1)*num/den` (join-crossover.rs:119-122).

Workloads (keys are the join/sort-major column):
  distinct   rows with 30% duplicate pairs, keys uniform on [0, 2^24) or Zipf(1.1)
  union      two sorted unique relations of `rows` pairs sharing 50%
  anti_join  the same two relations, left minus right
  join       probe = `rows` pairs with uniform or Zipf(1.1) keys on [0, 2^24);
             build = 2^24 pairs with uniform keys (about one match per probe)
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native, workload

KEY_SPACE = 1 << 24


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


def _dev(a: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32)).cuda()


def _temp(nbytes: int) -> torch.Tensor:
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda")


def percentile(sorted_values, num, den):
    return sorted_values[(len(sorted_values) - 1) * num // den]


class _Op:
    name = ""

    def algorithmic_bytes(self) -> float:
        raise NotImplementedError

    def launch(self, stream) -> None:
        raise NotImplementedError

    def result(self):
        raise NotImplementedError


class Distinct(_Op):
    name = "distinct"

    def __init__(self, rows: int, key_mode: int, seed: int):
        a, b = workload.distinct_input(seed, rows, 0.7, key_mode, KEY_SPACE)
        self.host = (a, b)
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
        return int(self.count.item())

    def algorithmic_bytes(self):
        return 8.0 * self.rows + 8.0 * self.out_rows()

    def result(self):
        u = self.out_rows()
        return (self.o0[:u].cpu().numpy().view(np.uint32), self.o1[:u].cpu().numpy().view(np.uint32))


class _SetOp(_Op):
    def __init__(self, rows: int, seed: int):
        (lf, ls), (rf, rs) = workload.overlapping_sets(seed, rows, 0.5)
        self.host = ((lf, ls), (rf, rs))
        self.rows = rows
        self.lf, self.ls, self.rf, self.rs = _dev(lf), _dev(ls), _dev(rf), _dev(rs)


class Union(_SetOp):
    name = "union"

    def __init__(self, rows: int, seed: int):
        super().__init__(rows, seed)
        self.o0 = torch.empty(2 * rows, dtype=torch.int32, device="cuda")
        self.o1 = torch.empty(2 * rows, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_union_u32_temporary_bytes(None, rows, None, rows, None, None, _p(self.count),
                                                              ctypes.byref(tb), None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_union_u32(_p(self.lf), _p(self.ls), self.rows, _p(self.rf), _p(self.rs),
                                                   self.rows, None, None, None, None, _p(self.count), _p(self.o0),
                                                   _p(self.o1), _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc

    def out_rows(self):
        return int(self.count.item())

    def algorithmic_bytes(self):
        return 16.0 * self.rows + 8.0 * self.out_rows()

    def result(self):
        u = self.out_rows()
        return (self.o0[:u].cpu().numpy().view(np.uint32), self.o1[:u].cpu().numpy().view(np.uint32))


class AntiJoin(_SetOp):
    name = "anti_join"

    def __init__(self, rows: int, seed: int):
        super().__init__(rows, seed)
        self.o0 = torch.empty(rows, dtype=torch.int32, device="cuda")
        self.o1 = torch.empty(rows, dtype=torch.int32, device="cuda")
        self.count = torch.zeros(1, dtype=torch.int32, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_anti_join_u32_temporary_bytes(None, None, _p(self.count), rows, ctypes.byref(tb),
                                                                  None)
        self.temp = _temp(tb.value)

    def launch(self, stream):
        rc = _native.hip().sparkalog_hip_anti_join_u32(_p(self.lf), _p(self.ls), self.rows, _p(self.rf), _p(self.rs),
                                                       self.rows, None, None, _p(self.count), _p(self.o0),
                                                       _p(self.o1), _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc

    def out_rows(self):
        return int(self.count.item()) & 0xFFFFFFFF

    def algorithmic_bytes(self):
        return 16.0 * self.rows + 8.0 * self.out_rows()

    def result(self):
        u = self.out_rows()
        return (self.o0[:u].cpu().numpy().view(np.uint32), self.o1[:u].cpu().numpy().view(np.uint32))


class Join(_Op):
    """probe (value, key) ⋈ build (key, value) on probe.1 = build.0 with
    output (probe.0, build.1) — the TC join plan."""

    name = "join"
    BUILD_ROWS = 1 << 24

    def __init__(self, rows: int, key_mode: int, seed: int):
        pk, pv = workload.pairs(seed, rows, key_mode, KEY_SPACE)
        bk, bv = workload.pairs(seed + 1, self.BUILD_ROWS, workload.KEY_UNIFORM, KEY_SPACE)
        self.host = ((pv, pk), (bk, bv))
        self.rows = rows
        # U32RangeIndex of build.0 (storage/src/lib.rs:357-393): rows by (value, row)
        order = np.argsort(bk, kind="stable").astype(np.uint32)
        sorted_keys = bk[order]
        keys, first = np.unique(sorted_keys, return_index=True)
        starts = np.append(first, bk.size).astype(np.uint32)
        self.index = (keys.astype(np.uint32), starts, order)
        self.l0, self.lk = _dev(pv), _dev(pk)
        self.r0, self.r1 = _dev(bk), _dev(bv)
        self.keys, self.starts, self.irows = _dev(self.index[0]), _dev(self.index[1]), _dev(self.index[2])
        self.counts = torch.empty(rows, dtype=torch.int64, device="cuda")
        self.offsets = torch.empty(rows, dtype=torch.int64, device="cuda")
        self.total = torch.zeros(1, dtype=torch.int64, device="cuda")
        tb = ctypes.c_size_t()
        _native.hip().sparkalog_hip_join_u32_temporary_bytes(_p(self.counts), _p(self.offsets), rows, ctypes.byref(tb),
                                                             None)
        self.temp = _temp(tb.value)
        self.o0 = self.o1 = None

    def launch(self, stream):
        lib = _native.hip()
        rc = lib.sparkalog_hip_join_u32_count(_p(self.lk), self.rows, _p(self.keys), _p(self.starts),
                                              self.keys.numel(), _p(self.counts), _p(self.offsets), _p(self.total),
                                              _p(self.temp), self.temp.numel(), stream)
        assert rc == 0, rc
        if self.o0 is None:
            # the reference synchronizes here to size the output (lib.rs:1608)
            torch.cuda.synchronize()
            m = int(self.total.item())
            self.o0 = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
            self.o1 = torch.empty(max(m, 1), dtype=torch.int32, device="cuda")
        rc = lib.sparkalog_hip_join_u32_emit(_p(self.lk), self.rows, _p(self.keys), _p(self.starts), _p(self.irows),
                                             self.keys.numel(), _p(self.l0), 0, _p(self.r1), 1, _p(self.offsets),
                                             _p(self.o0), _p(self.o1), stream)
        assert rc == 0, rc

    def out_rows(self):
        return int(self.total.item())

    def algorithmic_bytes(self):
        return 8.0 * self.rows + 8.0 * self.BUILD_ROWS + 8.0 * self.out_rows()

    def result(self):
        m = self.out_rows()
        return (self.o0[:m].cpu().numpy().view(np.uint32), self.o1[:m].cpu().numpy().view(np.uint32))


def make(name: str, rows: int, key_mode: int, seed: int = 0xC3) -> _Op:
    if name == "distinct":
        return Distinct(rows, key_mode, seed)
    if name == "union":
        return Union(rows, seed)
    if name == "anti_join":
        return AntiJoin(rows, seed)
    if name == "join":
        return Join(rows, key_mode, seed)
    raise ValueError(name)


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
