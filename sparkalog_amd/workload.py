"""Seeded synthetic workloads (include/sparkalog_workload.h) as numpy arrays.

The five configurations of the binary pipeline's benchmark:
  random_dag         - plain TC input: DAG of 1,000 nodes, 4,000 edges
  component_graph    - single-instance TC: 2^20 nodes in Zipf(1.5) components
  operator inputs    - pairs / distinct_input / overlapping_sets
  batch_graphs       - 4,096 independent instances of uneven size
  (served requests reuse component_graph with a per-request seed)
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native

C1_SEED = 0xC1
C2_SEED = 0xC2
C3_SEED = 0xC3
C4_SEED = 0xC4
C5_SEED = 0xC5


def _take(ptr, n, dtype) -> np.ndarray:
    lib = _native.workload()
    arr = np.ctypeslib.as_array(ptr, shape=(max(n, 1),))[:n].astype(dtype, copy=True)
    lib.sparkalog_wl_free(ctypes.cast(ptr, ctypes.c_void_p))
    return arr


def random_dag(seed: int = C1_SEED, nodes: int = 1000, edges: int = 4000) -> tuple[np.ndarray, np.ndarray]:
    lib = _native.workload()
    src = np.empty(edges, dtype=np.uint32)
    dst = np.empty(edges, dtype=np.uint32)
    rc = lib.sparkalog_wl_random_dag(seed, nodes, edges, src.ctypes.data_as(ctypes.c_void_p),
                                     dst.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise ValueError(f"cannot draw {edges} distinct edges over {nodes} nodes")
    return src, dst


def component_graph(seed: int = C2_SEED, nodes: int = 1 << 20, zipf_s: float = 1.5, size_min: int = 8,
                    size_max: int = 4096, mean_out_degree: float = 2.0):
    """Sorted unique edges (src, dst) and the component count."""
    lib = _native.workload()
    s = _native.u32p()
    d = _native.u32p()
    comps = ctypes.c_uint32(0)
    n = lib.sparkalog_wl_component_graph(seed, nodes, zipf_s, size_min, size_max, mean_out_degree,
                                         ctypes.byref(s), ctypes.byref(d), ctypes.byref(comps))
    if n == 0xFFFFFFFFFFFFFFFF:
        raise MemoryError("component graph generation failed")
    return _take(s, n, np.uint32), _take(d, n, np.uint32), comps.value


def batch_graphs(seed: int = C4_SEED, instances: int = 4096, n_min: int = 16, n_max: int = 16384,
                 lambda_min: float = 1.2, lambda_max: float = 3.0):
    """(src, dst, offsets[instances+1], node_counts[instances]); local ids."""
    lib = _native.workload()
    s = _native.u32p()
    d = _native.u32p()
    o = _native.u64p()
    c = _native.u32p()
    n = lib.sparkalog_wl_batch_graphs(seed, instances, n_min, n_max, lambda_min, lambda_max,
                                      ctypes.byref(s), ctypes.byref(d), ctypes.byref(o), ctypes.byref(c))
    if n == 0xFFFFFFFFFFFFFFFF:
        raise MemoryError("batch generation failed")
    return (_take(s, n, np.uint32), _take(d, n, np.uint32), _take(o, instances + 1, np.uint64),
            _take(c, instances, np.uint32))


def wide_labels(nodes: int, seed: int = 0xF011) -> np.ndarray:
    """Distinct u32 labels for node indexes 0..nodes-1 (nodes >= 2) spread over
    the whole u32 range: 0 and 0xFFFFFFFF always among them, the rest drawn
    uniformly, in random order. Relabelling a graph keeps its closure (as a
    set of pairs) up to the same relabelling."""
    if nodes < 2:
        raise ValueError("wide_labels needs at least two nodes")
    rng = np.random.default_rng(seed)
    vals = np.unique(rng.integers(1, 0xFFFFFFFF, size=nodes + nodes // 4 + 16, dtype=np.uint64))
    rng.shuffle(vals)
    if vals.size < nodes - 2:
        raise ValueError("too few distinct labels drawn")
    ids = np.concatenate([np.array([0, 0xFFFFFFFF], dtype=np.uint64), vals[: nodes - 2]]).astype(np.uint32)
    rng.shuffle(ids)
    return ids


def relabel(src, dst, labels) -> tuple[np.ndarray, np.ndarray]:
    """Edges (labels[src], labels[dst]) as sorted unique pairs."""
    key = np.unique((labels[src].astype(np.uint64) << np.uint64(32)) | labels[dst].astype(np.uint64))
    return (key >> np.uint64(32)).astype(np.uint32), (key & np.uint64(0xFFFFFFFF)).astype(np.uint32)


KEY_UNIFORM = 0
KEY_ZIPF = 1


def pairs(seed: int, rows: int, key_mode: int = KEY_UNIFORM, key_space: int = 1 << 24, zipf_s: float = 1.1):
    lib = _native.workload()
    a = np.empty(rows, dtype=np.uint32)
    b = np.empty(rows, dtype=np.uint32)
    lib.sparkalog_wl_pairs(seed, rows, key_mode, key_space, zipf_s, a.ctypes.data_as(ctypes.c_void_p),
                           b.ctypes.data_as(ctypes.c_void_p))
    return a, b


def distinct_input(seed: int, rows: int, unique_fraction: float = 0.7, key_mode: int = KEY_UNIFORM,
                   key_space: int = 1 << 24, zipf_s: float = 1.1):
    lib = _native.workload()
    a = np.empty(rows, dtype=np.uint32)
    b = np.empty(rows, dtype=np.uint32)
    lib.sparkalog_wl_distinct_input(seed, rows, unique_fraction, key_mode, key_space, zipf_s,
                                    a.ctypes.data_as(ctypes.c_void_p), b.ctypes.data_as(ctypes.c_void_p))
    return a, b


def overlapping_sets(seed: int, rows: int, overlap: float = 0.5, key_mode: int = KEY_UNIFORM,
                     key_space: int = 1 << 24, zipf_s: float = 1.1):
    lib = _native.workload()
    arrs = [np.empty(rows, dtype=np.uint32) for _ in range(4)]
    lib.sparkalog_wl_overlapping_sets(seed, rows, overlap, key_mode, key_space, zipf_s,
                                      *[x.ctypes.data_as(ctypes.c_void_p) for x in arrs])
    return (arrs[0], arrs[1]), (arrs[2], arrs[3])


TEXT_MIXED_SPACE, TEXT_COMMENTS, TEXT_CRLF, TEXT_SIGNS, TEXT_NO_FINAL_NEWLINE = 1, 2, 4, 8, 16


def edge_text(seed: int, rows: int, key_space: int = 0, style: int = 0):
    """(text as a uint8 array, first, second): `rows` facts in the loader's
    text format and the values written (the parse's ground truth)."""
    lib = _native.workload()
    first = np.empty(max(rows, 1), np.uint32)
    second = np.empty(max(rows, 1), np.uint32)
    ptr = ctypes.c_void_p()
    n = lib.sparkalog_wl_edge_text(seed, rows, key_space, style, ctypes.byref(ptr),
                                   first.ctypes.data_as(_native.u32p), second.ctypes.data_as(_native.u32p))
    if n == 0xFFFFFFFFFFFFFFFF:
        raise MemoryError("edge text generation failed")
    text = np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_uint8)), shape=(max(n, 1),))[:n].copy()
    lib.sparkalog_wl_free(ptr)
    return text, first[:rows], second[:rows]


def pair_digest(first, second) -> int:
    a = np.ascontiguousarray(first, dtype=np.uint32)
    b = np.ascontiguousarray(second, dtype=np.uint32)
    return _native.workload().sparkalog_wl_pair_digest(a.ctypes.data_as(ctypes.c_void_p),
                                                       b.ctypes.data_as(ctypes.c_void_p), a.size)


def ordered_pair_digest(first, second, chunk: int = 1 << 24) -> int:
    """Order-dependent digest of a pair sequence: the sum over rows i of
    mix64(((a_i << 32) | b_i) + (i + 1) * 0x9E3779B97F4A7C15) mod 2^64 (mix64
    as in pair_digest), so a row moved to another position changes it."""
    a = np.ascontiguousarray(first, dtype=np.uint32)
    b = np.ascontiguousarray(second, dtype=np.uint32)
    m1, m2 = np.uint64(0xBF58476D1CE4E5B9), np.uint64(0x94D049BB133111EB)
    golden = np.uint64(0x9E3779B97F4A7C15)
    total = np.uint64(0)
    with np.errstate(over="ignore"):
        for s in range(0, a.size, chunk):
            n = min(chunk, a.size - s)
            z = ((a[s:s + n].astype(np.uint64) << np.uint64(32)) | b[s:s + n].astype(np.uint64)) + \
                np.arange(s + 1, s + 1 + n, dtype=np.uint64) * golden
            z = (z ^ (z >> np.uint64(30))) * m1
            z = (z ^ (z >> np.uint64(27))) * m2
            total = total + np.sum(z ^ (z >> np.uint64(31)), dtype=np.uint64)
    return int(total)


def mix64(key: int) -> int:
    return _native.workload().sparkalog_wl_mix64(key)
