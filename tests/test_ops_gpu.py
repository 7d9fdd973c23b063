"""Operator parity on the GPU: every HIP operator, called through the C ABI,
against the oracle on the same seeded inputs (bit-exact, including order)."""
import numpy as np
import pytest
import torch

import oracle
from helpers import cols, pairs, random_pairs
from sparkalog_amd import execution as ex
from sparkalog_amd import plans, workload

pytestmark = pytest.mark.gpu


def dev(columns):
    return ex.Relation.from_host(*columns)


def host(rel):
    return rel.to_host()


def same(a, b):
    return np.array_equal(np.asarray(a[0]), np.asarray(b[0])) and np.array_equal(np.asarray(a[1]), np.asarray(b[1]))


# ---- known-answer vectors ----------------------------------------------------

def test_reference_vectors(golden, stream):
    for case in golden["distinct"]:
        out = ex.distinct_hip(dev(cols(case["input"])), plans.BinaryDistinct(), stream)
        assert out.pairs() == [tuple(p) for p in case["expected"]]
    for case in golden["anti_join"]:
        out = ex.anti_join_hip(dev(cols(case["left"])), dev(cols(case["right"])), plans.SortedBinaryAntiJoin(), stream)
        assert out.pairs() == [tuple(p) for p in case["expected"]]
    for case in golden["union"]:
        out = ex.union_hip(dev(cols(case["left"])), dev(cols(case["right"])), plans.SortedBinaryUnion(), stream)
        assert out.pairs() == [tuple(p) for p in case["expected"]]
    for case in golden["join"]:
        right = cols(case["right"])
        keys, starts, rows = oracle.range_index(right[case["plan"]["right_key"]])
        index = ex.RangeIndex.from_host(keys, starts, rows, len(right[0]))
        out = ex.join_hip(dev(cols(case["left"])), dev(right), index, plans.path_join_plan(), stream)
        assert out.pairs() == [tuple(p) for p in case["expected"]]


def test_filter_vectors(golden, stream):
    values = torch.from_numpy((np.arange(10000) % 97).astype(np.int32)).cuda()
    got = ex.filter_hip(values, plans.U32Predicate.ge(51), stream).cpu().numpy()
    assert list(got) == list(oracle.filter_rows(np.arange(10000) % 97, 5, 51))
    ten = torch.arange(10, dtype=torch.int32, device="cuda")
    for op, rows in golden["filter"][1]["expected"].items():
        assert list(ex.filter_hip(ten, plans.U32Predicate(int(op), 4), stream).cpu().numpy()) == rows
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    assert ex.filter_hip(empty, plans.U32Predicate.eq(0), stream).numel() == 0


def test_fill_and_add_one(stream):
    col = torch.zeros(257, dtype=torch.int32, device="cuda")
    ex.fill_mod_u32(col, 17, stream)
    assert list(col.cpu().numpy()) == [i % 17 for i in range(257)]
    ex.add_one_i32(col, stream)
    assert list(col.cpu().numpy()) == [i % 17 + 1 for i in range(257)]


@pytest.mark.parametrize("words", [1, 2, 7, 4096 * 4 + 3, 1_000_001])
def test_copy_u64(words, stream):
    src = torch.randint(-2**31, 2**31 - 1, (words * 2,), dtype=torch.int32, device="cuda")
    dst = torch.zeros(words * 2 + 2, dtype=torch.int32, device="cuda")
    ex.copy_u64(src, dst, stream)
    stream.synchronize()
    assert torch.equal(dst[:words * 2], src)
    assert int(dst[words * 2:].abs().sum()) == 0  # nothing past the copy


# ---- seeded parity -------------------------------------------------------------

DISTINCT_CASES = [
    (0, 1 << 32), (1, 1 << 32), (5, 3), (4095, 1 << 32), (4096, 7), (4097, 1 << 32),
    (100_000, 1000), (1_000_003, 1 << 32), (300_000, 1),
]


@pytest.mark.parametrize("n,hi", DISTINCT_CASES)
def test_distinct_random(n, hi, stream):
    rng = np.random.default_rng(n ^ hi)
    a = random_pairs(rng, n, hi)
    got = host(ex.distinct_hip(dev(a), plans.BinaryDistinct(), stream))
    assert same(got, oracle.distinct(*a, threads=4))


def test_distinct_swapped_columns(stream):
    rng = np.random.default_rng(9)
    a = random_pairs(rng, 50_000, 300)
    got = host(ex.distinct_hip(dev(a), plans.BinaryDistinct((1, 0)), stream))
    assert same(got, oracle.distinct(a[1], a[0]))


@pytest.mark.parametrize("key_mode", [workload.KEY_UNIFORM, workload.KEY_ZIPF])
def test_distinct_operator_input(key_mode, stream):
    a = workload.distinct_input(31 + key_mode, 1 << 21, key_mode=key_mode)
    got = host(ex.distinct_hip(dev(a), plans.BinaryDistinct(), stream))
    assert got[0].size == int(0.7 * (1 << 21) + 0.5)
    assert same(got, oracle.distinct(*a, threads=8))


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_ops_on_unaligned_column_views(offset, stream):
    # columns that start 4, 8 or 12 bytes into their allocation: whole tiles
    # are read as 16-byte vectors only when the columns are 16-byte aligned
    rng = np.random.default_rng(70 + offset)
    n = 3 * 4096 + 17
    a = random_pairs(rng, n + offset, 5000)
    full = dev(a)
    view = ex.Relation((full.columns[0][offset:], full.columns[1][offset:]))
    got = host(ex.distinct_hip(view, plans.BinaryDistinct(), stream))
    assert same(got, oracle.distinct(a[0][offset:], a[1][offset:]))
    left = oracle.distinct(*random_pairs(rng, n, 3000))
    right = oracle.distinct(*random_pairs(rng, n, 3000))
    # the relations one row into a larger allocation
    pad = lambda c: (np.concatenate([[7], c[0]]).astype(np.uint32), np.concatenate([[9], c[1]]).astype(np.uint32))
    lp, rp = dev(pad(left)), dev(pad(right))
    lv = ex.Relation((lp.columns[0][1:], lp.columns[1][1:]))
    rv = ex.Relation((rp.columns[0][1:], rp.columns[1][1:]))
    assert same(host(ex.union_hip(lv, rv, plans.SortedBinaryUnion(), stream)), oracle.union(left, right))
    assert same(host(ex.anti_join_hip(lv, rv, plans.SortedBinaryAntiJoin(), stream)), oracle.anti_join(left, right))


def _sorted_set(rng, n, hi):
    return oracle.distinct(*random_pairs(rng, n, hi))


SET_CASES = [(0, 0), (0, 100), (100, 0), (1, 1), (2048, 2048), (2049, 7), (5000, 300_000), (300_000, 5000),
             (1_000_000, 1_000_000), (10, 2_000_000)]


@pytest.mark.parametrize("nl,nr", SET_CASES)
def test_anti_join_random(nl, nr, stream):
    rng = np.random.default_rng(nl * 7 + nr)
    hi = max(8, int((nl + nr) ** 0.5) * 4)
    left, right = _sorted_set(rng, nl, hi), _sorted_set(rng, nr, hi)
    got = host(ex.anti_join_hip(dev(left), dev(right), plans.SortedBinaryAntiJoin(), stream))
    assert same(got, oracle.anti_join(left, right))


@pytest.mark.parametrize("nl,nr", SET_CASES)
def test_union_random(nl, nr, stream):
    rng = np.random.default_rng(nl * 13 + nr)
    hi = max(8, int((nl + nr) ** 0.5) * 4)
    left, right = _sorted_set(rng, nl, hi), _sorted_set(rng, nr, hi)
    got = host(ex.union_hip(dev(left), dev(right), plans.SortedBinaryUnion(), stream))
    assert same(got, oracle.union(left, right))


def test_union_duplicates_within_inputs(stream):
    rng = np.random.default_rng(77)
    # sorted inputs that keep duplicates: every adjacent duplicate must go
    a = random_pairs(rng, 200_000, 50)
    b = random_pairs(rng, 150_000, 50)
    ka = np.sort((a[0].astype(np.uint64) << 32) | a[1])
    kb = np.sort((b[0].astype(np.uint64) << 32) | b[1])
    left = ((ka >> 32).astype(np.uint32), (ka & 0xFFFFFFFF).astype(np.uint32))
    right = ((kb >> 32).astype(np.uint32), (kb & 0xFFFFFFFF).astype(np.uint32))
    got = host(ex.union_hip(dev(left), dev(right), plans.SortedBinaryUnion(), stream))
    assert same(got, oracle.union(left, right))


def test_anti_join_with_left_duplicates(stream):
    left = cols([(1, 1), (1, 1), (2, 2), (3, 3), (3, 3)])
    right = cols([(1, 1), (3, 4)])
    got = host(ex.anti_join_hip(dev(left), dev(right), plans.SortedBinaryAntiJoin(), stream))
    assert same(got, oracle.anti_join(left, right))


def test_operator_sets_config(stream):
    (lf, ls), (rf, rs) = workload.overlapping_sets(41, 1 << 20, 0.5)
    left, right = (lf, ls), (rf, rs)
    got = host(ex.anti_join_hip(dev(left), dev(right), plans.SortedBinaryAntiJoin(), stream))
    assert same(got, oracle.anti_join(left, right, threads=8))
    got = host(ex.union_hip(dev(left), dev(right), plans.SortedBinaryUnion(), stream))
    assert same(got, oracle.union(left, right, threads=8))


def _join_case(left, right, left_key, right_key, output, stream):
    keys, starts, rows = oracle.range_index(right[right_key])
    index = ex.RangeIndex.from_host(keys, starts, rows, len(right[0]))
    plan = plans.BinaryEqualityJoin(left_key, right_key, tuple(
        plans.JoinProjection(plans.JoinInput(side), col) for side, col in output))
    got = host(ex.join_hip(dev(left), dev(right), index, plan, stream))
    assert same(got, oracle.join(left, right, left_key, right_key, output, threads=4))


@pytest.mark.parametrize("seed", range(3))
def test_join_random(seed, stream):
    rng = np.random.default_rng(500 + seed)
    left = random_pairs(rng, 50_000, 5000)
    right = random_pairs(rng, 20_000, 5000)
    for left_key, right_key, output in [(1, 0, ((0, 0), (1, 1))), (0, 1, ((1, 0), (0, 1))), (1, 1, ((0, 1), (1, 0)))]:
        _join_case(left, right, left_key, right_key, output, stream)


def test_join_skewed_and_sparse(stream):
    rng = np.random.default_rng(600)
    # one hot key with 100k right matches, and long runs of unmatched left rows
    right = (np.concatenate([np.full(100_000, 7, np.uint32), rng.integers(1000, 2000, 5000, dtype=np.uint32)]),
             rng.integers(0, 1 << 31, 105_000, dtype=np.uint32))
    lk = rng.integers(3000, 4000, 50_000, dtype=np.uint32)       # unmatched
    lk[::10_000] = 7                                              # a few hot rows
    lk[5::7] = rng.integers(1000, 2000, lk[5::7].size, dtype=np.uint32)
    left = (rng.integers(0, 1 << 31, lk.size, dtype=np.uint32), lk)
    _join_case(left, right, 1, 0, ((0, 0), (1, 1)), stream)


def test_join_zipf_keys(stream):
    left = workload.pairs(71, 1 << 18, workload.KEY_ZIPF, key_space=1 << 16)
    right = workload.pairs(72, 1 << 16, workload.KEY_UNIFORM, key_space=1 << 16)
    _join_case((left[1], left[0]), right, 1, 0, ((0, 0), (1, 1)), stream)


@pytest.mark.parametrize("spread", ["uniform", "clustered"])
def test_join_long_key_column(spread, stream):
    """Index key columns far longer than the join's 2,048-entry LDS sample:
    keys uniform over 2^32 (the interpolated window holds the key) and keys
    in four tight clusters (the window misses and the whole bracket is
    searched); probes hit, miss by one, miss at random and sit at 0 and
    2^32 - 1."""
    rng = np.random.default_rng(700)
    n = 300_000
    if spread == "uniform":
        rk = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    else:
        rk = (rng.integers(0, 4, n, dtype=np.uint32) * np.uint32(1_000_000_000)
              + rng.integers(0, 400_000, n, dtype=np.uint32)).astype(np.uint32)
    right = (rk, rng.integers(0, 1 << 31, n, dtype=np.uint32))
    lk = np.concatenate([rk[rng.integers(0, n, 150_000)], (rk[:50_000].astype(np.uint64) + 1).astype(np.uint32),
                         rng.integers(0, 1 << 32, 50_000, dtype=np.uint64).astype(np.uint32),
                         np.array([0, 0xFFFFFFFF], np.uint32)])
    left = (rng.integers(0, 1 << 31, lk.size, dtype=np.uint32), lk)
    _join_case(left, right, 1, 0, ((0, 0), (1, 1)), stream)


@pytest.mark.parametrize("n", [1, 4095, 4096, 4097, 3_000_001])
def test_filter_random(n, stream):
    rng = np.random.default_rng(n)
    v = rng.integers(0, 1000, n, dtype=np.uint32)
    col = torch.from_numpy(v.view(np.int32)).cuda()
    for op in range(6):
        got = ex.filter_hip(col, plans.U32Predicate(op, 500), stream).cpu().numpy().view(np.uint32)
        assert np.array_equal(got, oracle.filter_rows(v, op, 500))


def test_row_limit_is_checked_on_the_host():
    with pytest.raises(ex.TooManyRows):
        ex._check_row_count(1 << 32)
