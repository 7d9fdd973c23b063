"""Operator placement policies (crates/sparkalog-execution/src/placement.rs).

The five policies keep the reference's fields and decision rule
(placement.rs:17-199): an operator runs on the GPU when one is available and
its row count reaches the threshold for where its input was produced;
without a GPU it runs on the parallel CPU path from its parallel threshold,
else serially.

The reference's thresholds (`MEASURED_GB10*`) were measured on GB10's unified
memory, where a CPU-produced input costs page migration. On MI355X an input
produced on the other side costs an explicit PCIe copy through pinned
memory, so the thresholds are re-derived here from the crossover harness
(`bench.py --mode crossover`, the methodology of the reference's
crates/sparkalog-cli/src/bin/*-crossover.rs): CPU serial and parallel times
of the oracle (the reference's CPU operators restated) against the HIP
operator with device-resident input, and with a CPU-produced input copied
to the device first. `derive` turns such a table into policies;
`MEASURED_MI355X` holds the values it gave on the MI355X box.
"""
from __future__ import annotations

from dataclasses import dataclass
from enum import Enum


class Placement(Enum):
    CPU_SERIAL = "cpu_serial"
    CPU_PARALLEL = "cpu_parallel"
    GPU = "gpu"


class InputProvenance(Enum):
    CPU = "cpu"
    GPU = "gpu"


def _place(rows: int, gpu_available: bool, gpu_min_rows: int, parallel_min_rows: int) -> Placement:
    if gpu_available and rows >= gpu_min_rows:
        return Placement.GPU
    if not gpu_available and rows >= parallel_min_rows:
        return Placement.CPU_PARALLEL
    return Placement.CPU_SERIAL


@dataclass(frozen=True)
class FilterPlacementPolicy:
    """placement.rs:24-54."""

    cpu_produced_gpu_min_rows: int
    gpu_produced_gpu_min_rows: int
    gpu_unavailable_parallel_min_rows: int

    def place(self, rows: int, input_provenance: InputProvenance, gpu_available: bool) -> Placement:
        gpu_min = (self.cpu_produced_gpu_min_rows if input_provenance == InputProvenance.CPU
                   else self.gpu_produced_gpu_min_rows)
        return _place(rows, gpu_available, gpu_min, self.gpu_unavailable_parallel_min_rows)


@dataclass(frozen=True)
class JoinPlacementPolicy:
    """placement.rs:62-88 (the delta is always produced by the previous
    operator, so there is no provenance)."""

    gpu_min_delta_rows: int
    gpu_unavailable_parallel_min_rows: int

    def place(self, delta_rows: int, gpu_available: bool) -> Placement:
        return _place(delta_rows, gpu_available, self.gpu_min_delta_rows, self.gpu_unavailable_parallel_min_rows)


@dataclass(frozen=True)
class DistinctPlacementPolicy:
    """placement.rs:98-128."""

    cpu_produced_gpu_min_rows: int
    gpu_produced_gpu_min_rows: int
    gpu_unavailable_parallel_min_rows: int

    def place(self, rows: int, input_provenance: InputProvenance, gpu_available: bool) -> Placement:
        gpu_min = (self.cpu_produced_gpu_min_rows if input_provenance == InputProvenance.CPU
                   else self.gpu_produced_gpu_min_rows)
        return _place(rows, gpu_available, gpu_min, self.gpu_unavailable_parallel_min_rows)


@dataclass(frozen=True)
class AntiJoinPlacementPolicy:
    """placement.rs:137-176: GPU and parallel thresholds both by provenance."""

    cpu_produced_gpu_min_rows: int
    gpu_produced_gpu_min_rows: int
    cpu_produced_parallel_min_rows: int
    gpu_produced_parallel_min_rows: int

    def place(self, left_rows: int, input_provenance: InputProvenance, gpu_available: bool) -> Placement:
        if input_provenance == InputProvenance.CPU:
            gpu_min, par_min = self.cpu_produced_gpu_min_rows, self.cpu_produced_parallel_min_rows
        else:
            gpu_min, par_min = self.gpu_produced_gpu_min_rows, self.gpu_produced_parallel_min_rows
        return _place(left_rows, gpu_available, gpu_min, par_min)


@dataclass(frozen=True)
class UnionPlacementPolicy:
    """placement.rs:185-218: thresholds on left + right rows."""

    cpu_produced_gpu_min_rows: int
    gpu_produced_gpu_min_rows: int
    gpu_unavailable_parallel_min_rows: int

    def place(self, left_rows: int, right_rows: int, input_provenance: InputProvenance,
              gpu_available: bool) -> Placement:
        rows = left_rows + right_rows
        gpu_min = (self.cpu_produced_gpu_min_rows if input_provenance == InputProvenance.CPU
                   else self.gpu_produced_gpu_min_rows)
        return _place(rows, gpu_available, gpu_min, self.gpu_unavailable_parallel_min_rows)


# The reference's GB10 values (placement.rs:34-38, :72-75, :108-112,
# :148-153, :195-199), kept for comparison.
MEASURED_GB10 = {
    "filter": FilterPlacementPolicy(131_072, 32_768, 8_388_608),
    "join": JoinPlacementPolicy(2_048, 2_048),
    "distinct": DistinctPlacementPolicy(32_768, 32_768, 131_072),
    "anti_join": AntiJoinPlacementPolicy(32_768, 16_384, 32_768, 262_144),
    "union": UnionPlacementPolicy(1_048_576, 1_048_576, 4_194_304),
}

NEVER = 1 << 62  # a threshold no row count reaches


def first_crossing(rows, faster) -> int:
    """The smallest measured row count from which `faster(i)` holds at every
    larger measured size (conservative, as the reference's thresholds are);
    NEVER when it does not hold at the largest."""
    best = NEVER
    for i in range(len(rows) - 1, -1, -1):
        if not faster(i):
            break
        best = rows[i]
    return best


def derive(table: dict) -> dict:
    """Policies from a crossover table: table[op] = {"rows": [...],
    "cpu_serial_ms": [...], "cpu_parallel_ms": [...], "gpu_ms": [...],
    "h2d_ms": [...], "d2h_ms": [...]} (medians per row count; rows is the
    policy's row measure: left rows for the anti-join, left + right for the
    union). A GPU-produced input run on the CPU pays d2h first; a
    CPU-produced input run on the GPU pays h2d first."""
    out = {}
    for op, t in table.items():
        rows = t["rows"]
        cpu = [min(a, b) for a, b in zip(t["cpu_serial_ms"], t["cpu_parallel_ms"])]
        gpu_prod = first_crossing(rows, lambda i: t["gpu_ms"][i] <= cpu[i] + t["d2h_ms"][i])
        cpu_prod = first_crossing(rows, lambda i: t["h2d_ms"][i] + t["gpu_ms"][i] <= cpu[i])
        par = first_crossing(rows, lambda i: t["cpu_parallel_ms"][i] <= t["cpu_serial_ms"][i])
        if op == "filter":
            out[op] = FilterPlacementPolicy(cpu_prod, gpu_prod, par)
        elif op == "join":
            out[op] = JoinPlacementPolicy(gpu_prod, par)
        elif op == "distinct":
            out[op] = DistinctPlacementPolicy(cpu_prod, gpu_prod, par)
        elif op == "anti_join":
            out[op] = AntiJoinPlacementPolicy(cpu_prod, gpu_prod, par, par)
        elif op == "union":
            out[op] = UnionPlacementPolicy(cpu_prod, gpu_prod, par)
    return out


# Derived by `derive` from `bench.py --mode crossover` on one MI355X and its
# host's CPU share (16 cores of an AMD EPYC), synthetic inputs of 2^8..2^24
# rows; where two runs differed (the filter's GPU threshold: 2^14 and 2^16),
# the larger value is kept.
MEASURED_MI355X = {
    "filter": FilterPlacementPolicy(65_536, 65_536, 4_194_304),
    "join": JoinPlacementPolicy(256, 4_096),
    "distinct": DistinctPlacementPolicy(16_384, 16_384, 1_048_576),
    "anti_join": AntiJoinPlacementPolicy(16_384, 16_384, 262_144, 262_144),
    "union": UnionPlacementPolicy(32_768, 8_192, 2_097_152),
}
