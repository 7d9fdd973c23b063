"""Operator API over device relations (the HIP backend of the reference's
crates/sparkalog-execution/src/lib.rs).

Each `*_hip` function is the counterpart of the reference's `*_cuda`
function followed by `.wait()`: it launches the gfx950 kernels through the C
ABI on a HIP stream, synchronizes, and returns the output trimmed to the
device-written row count. Relations are pairs of int32 device tensors holding
u32 bit patterns (the reference's two u32 columns). torch is used only to
allocate device memory and to name the stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _native
from .plans import (BinaryDistinct, BinaryEqualityJoin, JoinInput, SortedBinaryAntiJoin,
                    SortedBinaryUnion, U32Predicate)

U32_MAX = 0xFFFFFFFF


class ExecutionError(Exception):
    """execution::Error (crates/sparkalog-execution/src/lib.rs:171-188)."""


class HipError(ExecutionError):
    def __init__(self, operation: str, code: int):
        super().__init__(f"{operation} failed with HIP error {code}")
        self.operation = operation
        self.code = code


class TooManyRows(ExecutionError):
    def __init__(self, rows: int):
        super().__init__(f"{rows} rows cannot be represented by u32 row IDs")
        self.rows = rows


class MissingColumn(ExecutionError):
    def __init__(self, input_name: str, column: int):
        super().__init__(f"{input_name} relation has no column {column}")
        self.input = input_name
        self.column = column


class IndexSourceMismatch(ExecutionError):
    def __init__(self, relation_rows: int, index_rows: int):
        super().__init__(f"right relation has {relation_rows} rows but its index has {index_rows}")


def _check(operation: str, code: int) -> None:
    if code != 0:
        raise HipError(operation, code)


class DeviceWaitAbandoned(ExecutionError):
    """A kernel's bounded wait for another tile gave up: the ABI then writes
    an all-ones output count (no counterpart in the reference, whose CUB
    calls have no such wait)."""

    def __init__(self, operation: str):
        super().__init__(f"{operation}: a device tile wait gave up (all-ones output count)")


def _device_count(operation: str, t: "torch.Tensor", width: int) -> int:
    """The device-written output count of an operator (u32 or u64)."""
    mask = (1 << width) - 1
    value = int(t.item()) & mask
    if value == mask:
        raise DeviceWaitAbandoned(operation)
    return value


def _check_row_count(rows: int) -> None:
    # check_row_count, crates/sparkalog-execution/src/lib.rs:325-331
    if rows > U32_MAX:
        raise TooManyRows(rows)


class HipStream:
    """A HIP stream (default priority). Wraps torch's current stream unless
    another torch stream is given, so torch allocations and the kernels are
    ordered on the same queue."""

    def __init__(self, torch_stream: "torch.cuda.Stream | None" = None):
        self.torch_stream = torch_stream or torch.cuda.current_stream()
        self.raw = ctypes.c_void_p(self.torch_stream.cuda_stream)
        self.device = self.torch_stream.device

    def synchronize(self) -> None:
        _check("sparkalog_hip_stream_synchronize", _native.hip().sparkalog_hip_stream_synchronize(self.raw))


@dataclass
class Relation:
    """Two u32 columns on the device (int32 tensors holding u32 bit patterns)."""

    columns: tuple[torch.Tensor, torch.Tensor]

    def __len__(self) -> int:
        return int(self.columns[0].numel())

    def column(self, index: int, name: str = "input") -> torch.Tensor:
        if index not in (0, 1):
            raise MissingColumn(name, index)
        return self.columns[index]

    @staticmethod
    def from_host(first, second, device="cuda") -> "Relation":
        a = np.ascontiguousarray(np.asarray(first, dtype=np.uint32)).view(np.int32)
        b = np.ascontiguousarray(np.asarray(second, dtype=np.uint32)).view(np.int32)
        if a.shape != b.shape:
            raise ValueError("columns differ in length")
        return Relation((torch.from_numpy(a.copy()).to(device), torch.from_numpy(b.copy()).to(device)))

    def to_host(self) -> tuple[np.ndarray, np.ndarray]:
        return (self.columns[0].cpu().numpy().view(np.uint32), self.columns[1].cpu().numpy().view(np.uint32))

    def pairs(self) -> list[tuple[int, int]]:
        a, b = self.to_host()
        return list(zip(a.tolist(), b.tolist()))


def _ptr(t: "torch.Tensor | None"):
    return ctypes.c_void_p(t.data_ptr()) if t is not None and t.numel() > 0 else None


def _empty(n: int, dtype, device) -> torch.Tensor:
    return torch.empty(max(n, 1), dtype=dtype, device=device)


def _temp(nbytes: int, device) -> torch.Tensor:
    return torch.empty(max(nbytes, 256), dtype=torch.uint8, device=device)


def add_one_i32(column: torch.Tensor, stream: HipStream) -> None:
    """Smoke operator (crates/sparkalog-execution/src/lib.rs:305-323)."""
    _check("sparkalog_hip_add_one_i32",
           _native.hip().sparkalog_hip_add_one_i32(_ptr(column), column.numel(), stream.raw))
    stream.synchronize()


def copy_u64(src: torch.Tensor, dst: torch.Tensor, stream: HipStream) -> None:
    """Plain device copy of src's bytes into dst (asynchronous on `stream`)."""
    if src.nbytes % 8 or dst.nbytes < src.nbytes:
        raise ValueError("copy_u64: source must be whole u64 words and fit the destination")
    _check("sparkalog_hip_copy_u64", _native.hip().sparkalog_hip_copy_u64(_ptr(src), _ptr(dst), src.nbytes // 8,
                                                                          stream.raw))


def fill_mod_u32(column: torch.Tensor, modulus: int, stream: HipStream) -> None:
    """GPU producer of the crossover harnesses (lib.rs:1929-1948)."""
    _check("sparkalog_hip_fill_mod_u32",
           _native.hip().sparkalog_hip_fill_mod_u32(_ptr(column), column.numel(), modulus, stream.raw))
    stream.synchronize()


def filter_hip(column: torch.Tensor, predicate: U32Predicate, stream: HipStream) -> torch.Tensor:
    """filter_cuda + wait (lib.rs:1809-1875): ordered row ids of matches."""
    lib = _native.hip()
    rows = column.numel()
    _check_row_count(rows)
    dev = column.device
    out = _empty(rows, torch.int32, dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_filter_u32_temporary_bytes",
           lib.sparkalog_hip_filter_u32_temporary_bytes(None, _ptr(out), _ptr(count), rows, ctypes.byref(tb), stream.raw))
    temp = _temp(tb.value, dev)
    _check("sparkalog_hip_filter_u32",
           lib.sparkalog_hip_filter_u32(_ptr(column), rows, predicate.op, predicate.value, None, _ptr(out),
                                        _ptr(count), _ptr(temp), temp.numel(), stream.raw))
    stream.synchronize()
    return out[: _device_count("sparkalog_hip_filter_u32", count, 32)]


@dataclass
class RangeIndex:
    """U32RangeIndex (crates/sparkalog-storage/src/lib.rs:345-423) on the
    device: sorted unique keys, starts[K+1], rows sorted by (value, row)."""

    keys: torch.Tensor
    starts: torch.Tensor
    rows: torch.Tensor
    source_rows: int

    @staticmethod
    def from_host(keys, starts, rows, source_rows: int, device="cuda") -> "RangeIndex":
        def t(x):
            return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.uint32)).view(np.int32).copy()).to(device)
        return RangeIndex(t(keys), t(starts), t(rows), source_rows)

    @staticmethod
    def build(values: torch.Tensor, stream: "HipStream") -> "RangeIndex":
        """U32RangeIndex::build_slice (storage/src/lib.rs:357-393) on the
        device, over a u32 column (an int32 device tensor)."""
        lib = _native.hip()
        n = int(values.numel())
        _check_row_count(n)
        dev = values.device
        keys = _empty(n, torch.int32, dev)
        starts = _empty(n + 1, torch.int32, dev)
        rows = _empty(n, torch.int32, dev)
        k = torch.zeros(1, dtype=torch.int64, device=dev)
        tb = ctypes.c_size_t(0)
        _check("sparkalog_hip_range_index_u32_temporary_bytes",
               lib.sparkalog_hip_range_index_u32_temporary_bytes(n, ctypes.byref(tb)))
        temp = _temp(tb.value, dev)
        _check("sparkalog_hip_range_index_u32",
               lib.sparkalog_hip_range_index_u32(_ptr(values), n, _ptr(keys), _ptr(starts), _ptr(rows), _ptr(k),
                                                 _ptr(temp), temp.numel(), stream.raw))
        stream.synchronize()
        u = _device_count("sparkalog_hip_range_index_u32", k, 64)
        return RangeIndex(keys[:u], starts[:u + 1], rows[:n], n)


def join_hip(left: Relation, right: Relation, right_index: RangeIndex, plan: BinaryEqualityJoin,
             stream: HipStream) -> Relation:
    """join_cuda (lib.rs:1549-1648): count + scan, one sync for the exact
    output size, then emit."""
    lib = _native.hip()
    _check_row_count(len(left))
    _check_row_count(len(right))
    left_key = left.column(plan.left_key, "left")
    right.column(plan.right_key, "right")
    if len(right) != right_index.source_rows:
        raise IndexSourceMismatch(len(right), right_index.source_rows)
    proj = []
    for p in plan.output:
        src = left if p.input == JoinInput.LEFT else right
        proj.append((src.column(p.column, "left" if p.input == JoinInput.LEFT else "right"), int(p.input)))
    dev = left_key.device
    n = len(left)
    k = right_index.keys.numel()
    counts = _empty(n, torch.int64, dev)
    offsets = _empty(n, torch.int64, dev)
    total = torch.zeros(1, dtype=torch.int64, device=dev)
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_join_u32_temporary_bytes",
           lib.sparkalog_hip_join_u32_temporary_bytes(_ptr(counts), _ptr(offsets), n, ctypes.byref(tb), stream.raw))
    temp = _temp(tb.value, dev)
    _check("sparkalog_hip_join_u32_count",
           lib.sparkalog_hip_join_u32_count(_ptr(left_key), n, _ptr(right_index.keys), _ptr(right_index.starts), k,
                                            _ptr(counts), _ptr(offsets), _ptr(total), _ptr(temp), temp.numel(),
                                            stream.raw))
    stream.synchronize()
    m = _device_count("sparkalog_hip_join_u32_count", total, 64)
    out0 = _empty(m, torch.int32, dev)
    out1 = _empty(m, torch.int32, dev)
    if m:
        _check("sparkalog_hip_join_u32_emit",
               lib.sparkalog_hip_join_u32_emit(_ptr(left_key), n, _ptr(right_index.keys), _ptr(right_index.starts),
                                               _ptr(right_index.rows), k, _ptr(proj[0][0]), proj[0][1],
                                               _ptr(proj[1][0]), proj[1][1], _ptr(offsets), _ptr(out0), _ptr(out1),
                                               stream.raw))
        stream.synchronize()
    return Relation((out0[:m], out1[:m]))


def distinct_hip(input: Relation, plan: BinaryDistinct, stream: HipStream) -> Relation:
    """distinct_cuda + wait (lib.rs:465-531)."""
    lib = _native.hip()
    _check_row_count(len(input))
    a = input.column(plan.columns[0])
    b = input.column(plan.columns[1])
    n = len(input)
    dev = a.device
    packed = _empty(n, torch.int64, dev)
    scratch = _empty(n, torch.int64, dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    out0 = _empty(n, torch.int32, dev)
    out1 = _empty(n, torch.int32, dev)
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_distinct_u32_temporary_bytes",
           lib.sparkalog_hip_distinct_u32_temporary_bytes(_ptr(packed), _ptr(scratch), _ptr(count), n,
                                                          ctypes.byref(tb), stream.raw))
    temp = _temp(tb.value, dev)
    _check("sparkalog_hip_distinct_u32",
           lib.sparkalog_hip_distinct_u32(_ptr(a), _ptr(b), n, _ptr(packed), _ptr(scratch), _ptr(count), _ptr(out0),
                                          _ptr(out1), _ptr(temp), temp.numel(), stream.raw))
    stream.synchronize()
    u = _device_count("sparkalog_hip_distinct_u32", count, 64)
    return Relation((out0[:u], out1[:u]))


def anti_join_hip(left: Relation, right: Relation, plan: SortedBinaryAntiJoin, stream: HipStream) -> Relation:
    """anti_join_cuda + wait (lib.rs:814-886)."""
    lib = _native.hip()
    _check_row_count(len(left))
    _check_row_count(len(right))
    lf, ls = left.column(plan.left[0], "left"), left.column(plan.left[1], "left")
    rf, rs = right.column(plan.right[0], "right"), right.column(plan.right[1], "right")
    n = len(left)
    dev = lf.device
    out0 = _empty(n, torch.int32, dev)
    out1 = _empty(n, torch.int32, dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_anti_join_u32_temporary_bytes",
           lib.sparkalog_hip_anti_join_u32_temporary_bytes(None, None, _ptr(count), n, ctypes.byref(tb), stream.raw))
    temp = _temp(tb.value, dev)
    _check("sparkalog_hip_anti_join_u32",
           lib.sparkalog_hip_anti_join_u32(_ptr(lf), _ptr(ls), n, _ptr(rf), _ptr(rs), len(right), None, None,
                                           _ptr(count), _ptr(out0), _ptr(out1), _ptr(temp), temp.numel(),
                                           stream.raw))
    stream.synchronize()
    o = _device_count("sparkalog_hip_anti_join_u32", count, 32)
    return Relation((out0[:o], out1[:o]))


def union_hip(left: Relation, right: Relation, plan: SortedBinaryUnion, stream: HipStream) -> Relation:
    """union_cuda + wait (lib.rs:1098-1177)."""
    lib = _native.hip()
    _check_row_count(len(left))
    _check_row_count(len(right))
    lf, ls = left.column(plan.left[0], "left"), left.column(plan.left[1], "left")
    rf, rs = right.column(plan.right[0], "right"), right.column(plan.right[1], "right")
    n = len(left) + len(right)
    dev = lf.device
    out0 = _empty(n, torch.int32, dev)
    out1 = _empty(n, torch.int32, dev)
    count = torch.zeros(1, dtype=torch.int64, device=dev)
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_union_u32_temporary_bytes",
           lib.sparkalog_hip_union_u32_temporary_bytes(None, len(left), None, len(right), None, None, _ptr(count),
                                                       ctypes.byref(tb), stream.raw))
    temp = _temp(tb.value, dev)
    _check("sparkalog_hip_union_u32",
           lib.sparkalog_hip_union_u32(_ptr(lf), _ptr(ls), len(left), _ptr(rf), _ptr(rs), len(right), None, None,
                                       None, None, _ptr(count), _ptr(out0), _ptr(out1), _ptr(temp), temp.numel(),
                                       stream.raw))
    stream.synchronize()
    u = _device_count("sparkalog_hip_union_u32", count, 64)
    return Relation((out0[:u], out1[:u]))
