"""sparkalog_hip_join_u32_count_no_alloc (the join count that allocates
nothing) against sparkalog_hip_join_u32_count on the same inputs, through the
C ABI: two calls in a row, the second with a smaller input, bit-exact."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
from sparkalog_amd import _native

pytestmark = pytest.mark.gpu


def _t(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.uint32)).view(np.int32).copy()).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _inputs(seed, left_rows, right_rows, key_range):
    rng = np.random.default_rng(seed)
    right_keys = rng.integers(0, key_range, size=right_rows, dtype=np.uint32)
    keys, starts, _ = oracle.range_index(right_keys)
    left = rng.integers(0, key_range + key_range // 4, size=left_rows, dtype=np.uint32)
    return left, keys, starts


def _count(lib, stream, left, keys, starts, no_alloc):
    n = left.size
    lk, ik, ist = _t(left), _t(keys), _t(starts)
    counts = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    offsets = torch.full((n,), -1, dtype=torch.int64, device="cuda")
    total = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    if no_alloc:
        rc = lib.sparkalog_hip_join_u32_count_no_alloc(_p(lk), n, _p(ik), _p(ist), keys.size, _p(counts),
                                                       _p(offsets), _p(total), stream.raw)
    else:
        tb = ctypes.c_size_t(0)
        assert lib.sparkalog_hip_join_u32_temporary_bytes(_p(counts), _p(offsets), n, ctypes.byref(tb),
                                                          stream.raw) == 0
        temp = torch.empty(max(tb.value, 256), dtype=torch.uint8, device="cuda")
        rc = lib.sparkalog_hip_join_u32_count(_p(lk), n, _p(ik), _p(ist), keys.size, _p(counts), _p(offsets),
                                              _p(total), _p(temp), temp.numel(), stream.raw)
    assert rc == 0
    stream.synchronize()
    return counts.cpu().numpy(), offsets.cpu().numpy(), int(total.item())


def test_no_alloc_join_count_matches_join_count(stream):
    lib = _native.hip()
    # 3,000,000 left rows (1,465 look-back tiles), then 50,000 (25 tiles): the
    # second call must not see the first call's status words
    for seed, left_rows, right_rows, key_range in [(0xA1, 3_000_000, 400_000, 1 << 20),
                                                    (0xA2, 50_000, 3_000, 5_000)]:
        print(f"no_alloc join count: seed {seed:#x}, {left_rows} left rows", flush=True)
        left, keys, starts = _inputs(seed, left_rows, right_rows, key_range)
        got = _count(lib, stream, left, keys, starts, no_alloc=True)
        ref = _count(lib, stream, left, keys, starts, no_alloc=False)
        assert np.array_equal(got[0], ref[0])
        assert np.array_equal(got[1], ref[1])
        assert got[2] == ref[2]
        # and against the matches counted on the host
        pos = np.searchsorted(keys, left)
        hit = (pos < keys.size) & (keys[np.minimum(pos, keys.size - 1)] == left)
        expect = np.where(hit, starts[np.minimum(pos + 1, keys.size)] - starts[np.minimum(pos, keys.size - 1)], 0)
        assert np.array_equal(got[0], expect.astype(np.int64))
        assert got[2] == int(expect.sum())


def test_no_alloc_join_count_empty_and_oversized(stream):
    lib = _native.hip()
    total = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    assert lib.sparkalog_hip_join_u32_count_no_alloc(None, 0, None, None, 0, None, None, _p(total), stream.raw) == 0
    stream.synchronize()
    assert int(total.item()) == 0
    # past the static scratch: refused before any device work
    assert lib.sparkalog_hip_join_u32_count_no_alloc(None, 268369921, None, None, 0, None, None, _p(total),
                                                     stream.raw) != 0
