"""Device range index (sparkalog_hip_range_index_u32) against the oracle's
restatement of U32RangeIndex::build_slice (storage/src/lib.rs:357-393),
exact arrays, and the ABI join driven by it against the oracle's join."""
import numpy as np
import pytest
import torch

import oracle
from helpers import random_pairs
from sparkalog_amd import execution as ex
from sparkalog_amd import plans

pytestmark = pytest.mark.gpu


def _col(values):
    return torch.from_numpy(np.ascontiguousarray(values, dtype=np.uint32).view(np.int32)).cuda()


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


def _same_index(values, stream):
    idx = ex.RangeIndex.build(_col(values), stream)
    keys, starts, rows = oracle.range_index(values)
    assert np.array_equal(_u32(idx.keys), keys)
    assert np.array_equal(_u32(idx.starts), starts)
    assert np.array_equal(_u32(idx.rows), rows)
    assert idx.source_rows == len(values)


def test_index_vectors(golden, stream):
    for case in golden["range_index"]:
        _same_index(np.asarray(case["values"], np.uint32), stream)


@pytest.mark.parametrize("n,space", [(0, 1), (1, 1), (4095, 10), (4096, 4096), (4097, 1 << 32),
                                     (100_000, 7), (1 << 20, 1 << 32), (1 << 22, 1 << 16)])
def test_index_random(n, space, stream):
    rng = np.random.default_rng(n + 17)
    values = rng.integers(0, space, n, dtype=np.uint64).astype(np.uint32)
    _same_index(values, stream)


def test_index_extremes(stream):
    # 0 and 0xFFFFFFFF, one value repeated across many tiles
    values = np.array([0xFFFFFFFF, 0, 0xFFFFFFFF, 5] + [9] * 20_000 + [0], np.uint32)
    _same_index(values, stream)


@pytest.mark.parametrize("seed", range(2))
def test_join_with_device_index(seed, stream):
    rng = np.random.default_rng(900 + seed)
    left = random_pairs(rng, 80_000, 6000)
    right = random_pairs(rng, 30_000, 6000)
    for left_key, right_key, output in [(1, 0, ((0, 0), (1, 1))), (0, 1, ((1, 0), (0, 1)))]:
        r = ex.Relation.from_host(*right)
        index = ex.RangeIndex.build(r.column(right_key), stream)
        plan = plans.BinaryEqualityJoin(left_key, right_key, tuple(
            plans.JoinProjection(plans.JoinInput(side), col) for side, col in output))
        got = ex.join_hip(ex.Relation.from_host(*left), r, index, plan, stream).to_host()
        exp = oracle.join(left, right, left_key, right_key, output, threads=4)
        assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1])
