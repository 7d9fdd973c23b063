"""The C ABI library loads on the host and exports every symbol that
include/*.h declares (no GPU calls)."""
import ctypes
import re
from pathlib import Path

import pytest

from sparkalog_amd import _native

ROOT = Path(__file__).resolve().parents[1]


def declared(header: Path) -> list[str]:
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b((?:sparkalog_hip|sparkalog_wl)_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,loader", [("sparkalog_hip.h", _native.hip), ("sparkalog_workload.h", _native.workload)])
def test_every_declared_symbol_is_exported(header, loader):
    lib = loader()
    names = declared(ROOT / "include" / header)
    assert names, header
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_the_header():
    names = set(declared(ROOT / "include" / "sparkalog_hip.h"))
    assert names == set(_native.hip_exported_symbols())
    assert set(declared(ROOT / "include" / "sparkalog_workload.h")) == set(_native.workload_exported_symbols())


def test_reference_operator_set_is_mirrored():
    # the 14 extern "C" functions of column_ops.cu:231-775, one-for-one
    reference = ["add_one_i32", "stream_synchronize", "filter_u32_temporary_bytes", "filter_u32", "fill_mod_u32",
                 "join_u32_temporary_bytes", "join_u32_count", "join_u32_emit", "distinct_u32_temporary_bytes",
                 "distinct_u32", "anti_join_u32_temporary_bytes", "anti_join_u32", "union_u32_temporary_bytes",
                 "union_u32"]
    lib = _native.hip()
    for name in reference:
        assert hasattr(lib, "sparkalog_hip_" + name)


def test_host_side_argument_validation_needs_no_gpu():
    lib = _native.hip()
    size = ctypes.c_size_t(0)
    # null count -> invalid value (column_ops.cu:258-260), before any device work
    assert lib.sparkalog_hip_filter_u32_temporary_bytes(None, None, None, 10, ctypes.byref(size), None) != 0
    assert lib.sparkalog_hip_distinct_u32(None, None, 4, None, None, None, None, None, None, 0, None) != 0
    assert lib.sparkalog_hip_fill_mod_u32(None, 4, 0, None) != 0
    # the no-alloc join count refuses inputs past its static scratch first
    dummy_total = ctypes.c_void_p(16)
    assert lib.sparkalog_hip_join_u32_count_no_alloc(None, 268369921, None, None, 0, None, None, dummy_total,
                                                     None) != 0
    assert lib.sparkalog_hip_join_u32_count_no_alloc(None, 8, None, None, 0, None, None, None, None) != 0
    # temporary-size queries are pure host arithmetic
    dummy = ctypes.c_void_p(16)
    assert lib.sparkalog_hip_union_u32_temporary_bytes(None, 100, None, 50, None, None, dummy, ctypes.byref(size), None) == 0
    assert size.value > 0


def test_all_ones_count_is_a_device_error():
    # the ABI's report of an abandoned device wait reaches the caller as an
    # exception, never as a row count
    import torch

    from sparkalog_amd import execution as ex

    assert ex._device_count("op", torch.tensor([5], dtype=torch.int64), 64) == 5
    assert ex._device_count("op", torch.tensor([7], dtype=torch.int32), 32) == 7
    with pytest.raises(ex.DeviceWaitAbandoned):
        ex._device_count("op", torch.tensor([-1], dtype=torch.int64), 64)
    with pytest.raises(ex.DeviceWaitAbandoned):
        ex._device_count("op", torch.tensor([-1], dtype=torch.int32), 32)


def test_status_strings():
    lib = _native.hip()
    assert lib.sparkalog_hip_status_string(4).decode().startswith("recursive SCC did not converge")
