"""The operator microbench configuration on the GPU, through the C ABI:
BinaryDistinct (30% duplicate pairs), SortedBinaryUnion and
SortedBinaryAntiJoin (two sorted unique relations sharing 50%) and
BinaryEqualityJoin (probe rows against a 2^24-row build), keys uniform on
[0, 2^24) or Zipf(1.1). At 2^24 rows the outputs are compared array for
array, in order, with the oracle; at 2^26 and 2^28 rows with the fixtures
that tests/golden/make_op_fixtures.py wrote from the oracle (row count,
order-independent digest, strict sortedness, and for the join an
order-dependent digest), after the input's digest pins the generator."""
import ctypes
import json
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN
from helpers import operator_input_digest, oracle_operator, ordered_digest, strictly_sorted
from sparkalog_amd import opbench, workload

pytestmark = pytest.mark.gpu

MODES = {"uniform": workload.KEY_UNIFORM, "zipf1.1": workload.KEY_ZIPF}
THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def op_expected():
    return json.loads((GOLDEN / "op_expected.json").read_text())


_HOST_CACHE: dict = {}


def _host(name, k, mode):
    # union and anti-join run on the same two relations
    group = "sets" if name in ("union", "anti_join") else name
    key = (group, k, mode)
    if key not in _HOST_CACHE:
        _HOST_CACHE.clear()
        _HOST_CACHE[key] = opbench.host_inputs(name, 1 << k, MODES[mode])
    return _HOST_CACHE[key]


def _run(name, host):
    op = opbench.make(name, 0, 0, host=host)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    op.launch(stream)
    torch.cuda.synchronize()
    got = op.result()
    del op
    torch.cuda.empty_cache()
    return got


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", opbench.OPERATORS)
def test_operator_config_2p24(name, mode, stream):
    print(f"operator config: {name}, 2^24 rows, {mode} keys", flush=True)
    host = _host(name, 24, mode)
    got = _run(name, host)
    exp = oracle_operator(name, host, THREADS)
    assert got[0].size == exp[0].size
    assert np.array_equal(got[0], exp[0]) and np.array_equal(got[1], exp[1])


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("name", opbench.OPERATORS)
@pytest.mark.parametrize("k", [26, 28])
def test_operator_config_large(k, name, mode, op_expected, stream):
    print(f"operator config: {name}, 2^{k} rows, {mode} keys", flush=True)
    e = op_expected[f"{name}_2p{k}_{mode}"]
    host = _host(name, k, mode)
    assert operator_input_digest(name, host) == e["input_digest"]
    got = _run(name, host)
    assert got[0].size == e["out_rows"]
    assert workload.pair_digest(*got) == e["digest"]
    if name == "join":
        assert ordered_digest(*got) == e["ordered_digest"]
    else:
        assert strictly_sorted(*got)
