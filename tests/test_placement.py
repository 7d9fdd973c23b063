"""Placement policies (crates/sparkalog-execution/src/placement.rs): the
reference's own tests (placement.rs:220-411) on its GB10 values, the
derivation of thresholds from a crossover table, and the MI355X values."""
import pytest

from sparkalog_amd import placement as pl
from sparkalog_amd.placement import InputProvenance as P, Placement as Pl


def test_filter_policy_accounts_for_provenance():
    f = pl.MEASURED_GB10["filter"]
    assert f.place(32_768, P.CPU, True) == Pl.CPU_SERIAL
    assert f.place(32_768, P.GPU, True) == Pl.GPU
    assert f.place(131_072, P.CPU, True) == Pl.GPU
    # parallel only as the GPU-unavailable fallback
    assert f.place(8_388_607, P.CPU, False) == Pl.CPU_SERIAL
    assert f.place(8_388_608, P.CPU, False) == Pl.CPU_PARALLEL
    assert f.place(8_388_608, P.CPU, True) == Pl.GPU


def test_join_policy():
    j = pl.MEASURED_GB10["join"]
    assert j.place(512, True) == Pl.CPU_SERIAL
    assert j.place(2_048, True) == Pl.GPU
    assert j.place(2_048, False) == Pl.CPU_PARALLEL


def test_distinct_policy():
    d = pl.MEASURED_GB10["distinct"]
    for prov in (P.CPU, P.GPU):
        assert d.place(7_424, prov, True) == Pl.CPU_SERIAL
        assert d.place(32_768, prov, True) == Pl.GPU
        assert d.place(131_072, prov, False) == Pl.CPU_PARALLEL


def test_anti_join_policy():
    a = pl.MEASURED_GB10["anti_join"]
    assert a.place(16_384, P.CPU, True) == Pl.CPU_SERIAL
    assert a.place(16_384, P.GPU, True) == Pl.GPU
    assert a.place(32_768, P.CPU, False) == Pl.CPU_PARALLEL
    assert a.place(131_072, P.GPU, False) == Pl.CPU_SERIAL
    assert a.place(262_144, P.GPU, False) == Pl.CPU_PARALLEL


def test_union_policy_uses_total_rows():
    u = pl.MEASURED_GB10["union"]
    assert u.place(1_000_000, 40_000, P.CPU, True) == Pl.CPU_SERIAL
    assert u.place(1_000_000, 50_000, P.GPU, True) == Pl.GPU
    assert u.place(2_000_000, 2_194_304, P.CPU, False) == Pl.CPU_PARALLEL


def test_first_crossing_is_conservative():
    rows = [1, 2, 4, 8, 16]
    # faster at 2, slower at 4, faster from 8 on: the threshold is 8
    assert pl.first_crossing(rows, lambda i: i != 2 and i != 0) == 8
    assert pl.first_crossing(rows, lambda i: False) == pl.NEVER
    assert pl.first_crossing(rows, lambda i: True) == 1


def test_derive_from_a_crossover_table():
    rows = [1 << 10, 1 << 12, 1 << 14, 1 << 16]
    table = {
        "distinct": {"rows": rows, "cpu_serial_ms": [0.05, 0.3, 1.5, 8.0], "cpu_parallel_ms": [0.2, 0.3, 0.6, 2.0],
                     "gpu_ms": [0.1, 0.12, 0.15, 0.2], "h2d_ms": [0.02, 0.05, 0.5, 2.5],
                     "d2h_ms": [0.02, 0.05, 0.5, 2.5]},
        "join": {"rows": rows, "cpu_serial_ms": [0.01, 0.05, 0.2, 1.0], "cpu_parallel_ms": [0.1, 0.1, 0.15, 0.3],
                 "gpu_ms": [0.1, 0.1, 0.1, 0.1], "h2d_ms": [0, 0, 0, 0], "d2h_ms": [0, 0, 0, 0]},
    }
    d = pl.derive(table)
    # GPU-produced: 0.1 > 0.05 + 0.02 at 2^10, then the GPU wins
    assert d["distinct"].gpu_produced_gpu_min_rows == 1 << 12
    # CPU-produced: h2d + gpu = 0.17 > 0.3? no: 0.17 <= 0.3 at 2^12; 0.65 > 0.6 at 2^14; 2.7 > 2.0 at 2^16
    assert d["distinct"].cpu_produced_gpu_min_rows == pl.NEVER
    assert d["distinct"].gpu_unavailable_parallel_min_rows == 1 << 12
    assert d["join"].gpu_min_delta_rows == 1 << 14
    assert d["join"].gpu_unavailable_parallel_min_rows == 1 << 14


@pytest.mark.skipif(not pl.MEASURED_MI355X, reason="MI355X values not recorded")
def test_mi355x_values_route_large_inputs_to_the_gpu():
    m = pl.MEASURED_MI355X
    assert set(m) == {"filter", "join", "distinct", "anti_join", "union"}
    assert m["distinct"].place(1 << 24, P.GPU, True) == Pl.GPU
    assert m["join"].place(1 << 24, True) == Pl.GPU
    assert m["union"].place(1 << 23, 1 << 23, P.GPU, True) == Pl.GPU
    for op in ("distinct", "anti_join", "union", "filter"):
        pol = m[op]
        # an input that must cross PCIe first never moves to the GPU sooner
        assert pol.cpu_produced_gpu_min_rows >= pol.gpu_produced_gpu_min_rows
