"""Shared helpers of the test suite (independent set formulations)."""
import numpy as np


def cols(pairs):
    """[(a, b), ...] -> (uint32 array, uint32 array)."""
    if len(pairs) == 0:
        return np.empty(0, np.uint32), np.empty(0, np.uint32)
    a = np.array([p[0] for p in pairs], dtype=np.uint32)
    b = np.array([p[1] for p in pairs], dtype=np.uint32)
    return a, b


def pairs(columns):
    a, b = columns
    return list(zip(np.asarray(a).tolist(), np.asarray(b).tolist()))


def packed(columns):
    a, b = columns
    return (np.asarray(a, dtype=np.uint64) << np.uint64(32)) | np.asarray(b, dtype=np.uint64)


def chain(n, start=0):
    """Edges i -> i+1 over n vertices (tc-fixpoint.rs:120-125)."""
    a = np.arange(start, start + n - 1, dtype=np.uint32)
    return a, a + np.uint32(1)


def cycle(n, start=0):
    a = np.arange(start, start + n, dtype=np.uint32)
    return a, np.roll(a, -1)


def random_pairs(rng, n, hi):
    return rng.integers(0, hi, size=n, dtype=np.uint32), rng.integers(0, hi, size=n, dtype=np.uint32)


def sorted_unique(columns):
    return tuple(np.asarray(x, dtype=np.uint32) for x in zip(*sorted(set(pairs(columns))))) \
        if len(columns[0]) else (np.empty(0, np.uint32), np.empty(0, np.uint32))


def ordered_digest(first, second):
    """Order-dependent digest of a pair sequence (workload.ordered_pair_digest)."""
    from sparkalog_amd import workload

    return workload.ordered_pair_digest(first, second)


def strictly_sorted(first, second) -> bool:
    k = packed((first, second))
    return bool(np.all(k[1:] > k[:-1]))


def oracle_operator(name, host, threads):
    """The oracle's result of one operator microbench case
    (sparkalog_amd.opbench.host_inputs)."""
    import oracle

    if name == "distinct":
        return oracle.distinct(*host["input"], threads=threads)
    if name == "union":
        return oracle.union(host["left"], host["right"], threads=threads)
    if name == "anti_join":
        return oracle.anti_join(host["left"], host["right"], threads=threads)
    return oracle.join(host["left"], host["right"], 1, 0, ((0, 0), (1, 1)), threads=threads)


def operator_input_digest(name, host):
    from sparkalog_amd import workload

    if name == "distinct":
        return int(workload.pair_digest(*host["input"]))
    return int(sum(workload.pair_digest(*p) for p in (host["left"], host["right"])) % (1 << 64))
