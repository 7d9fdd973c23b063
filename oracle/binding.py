"""ctypes binding of the oracle (test infrastructure only; see oracle.h).

Builds oracle/_build/liboracle.so with `make -C oracle` when it is missing
and the reference-free C sources are present.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_int, c_uint32, c_uint64, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"

__all__ = ["Rule", "Summary", "BfsResult", "lib", "build", "distinct", "anti_join", "union", "range_index",
           "join", "filter_rows", "tc_bfs", "general_tc", "Store", "rule_from_plan", "ORACLE_ITERATION_LIMIT",
           "load_binary_u32", "LOAD_ERRORS"]

ORACLE_ITERATION_LIMIT = 4


class Rule(ctypes.Structure):
    _fields_ = [
        ("target", c_uint32), ("delta_input", c_uint32), ("right_input", c_uint32),
        ("left_key", c_uint32), ("right_key", c_uint32),
        ("out_input", c_uint32 * 2), ("out_column", c_uint32 * 2), ("distinct_columns", c_uint32 * 2),
        ("anti_left", c_uint32 * 2), ("anti_right", c_uint32 * 2),
        ("union_left", c_uint32 * 2), ("union_right", c_uint32 * 2),
    ]


class Summary(ctypes.Structure):
    _fields_ = [("iterations", c_uint64), ("total_new_rows", c_uint64), ("delta_rows_left", c_uint64)]


class BfsResult(ctypes.Structure):
    _fields_ = [("pairs", c_uint64), ("digest", c_uint64), ("max_dist", c_uint64)]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = ctypes.CDLL(str(LIB))
        u32 = c_void_p
        L.oracle_sort_u64.argtypes = [c_void_p, c_uint64, c_int]
        L.oracle_distinct.restype = c_uint64
        L.oracle_distinct.argtypes = [u32, u32, c_uint64, u32, u32, c_int]
        L.oracle_anti_join.restype = c_uint64
        L.oracle_anti_join.argtypes = [u32, u32, c_uint64, u32, u32, c_uint64, u32, u32, c_int]
        L.oracle_union.restype = c_uint64
        L.oracle_union.argtypes = [u32, u32, c_uint64, u32, u32, c_uint64, u32, u32, c_int]
        L.oracle_range_index.restype = c_uint64
        L.oracle_range_index.argtypes = [u32, c_uint64, u32, u32, u32]
        L.oracle_join_count.restype = c_uint64
        L.oracle_join_count.argtypes = [u32, c_uint64, u32, u32, c_uint64, c_int]
        L.oracle_join.restype = c_uint64
        L.oracle_join.argtypes = [u32, c_uint64, u32, u32, u32, c_uint64, u32, c_uint32, u32, c_uint32, u32, u32, c_int]
        L.oracle_filter.restype = c_uint64
        L.oracle_filter.argtypes = [u32, c_uint64, c_uint32, c_uint32, u32, c_int]
        L.oracle_store_new.restype = c_void_p
        L.oracle_store_free.argtypes = [c_void_p]
        L.oracle_store_insert_static.restype = c_int
        L.oracle_store_insert_static.argtypes = [c_void_p, c_uint32, u32, u32, c_uint64]
        L.oracle_store_insert_recursive.restype = c_int
        L.oracle_store_insert_recursive.argtypes = [c_void_p, c_uint32, u32, u32, c_uint64, u32, u32, c_uint64]
        L.oracle_store_get.restype = c_uint64
        L.oracle_store_get.argtypes = [c_void_p, c_uint32, c_int, u32, u32]
        L.oracle_run.restype = c_int
        L.oracle_run.argtypes = [c_void_p, POINTER(c_uint32), c_uint32, POINTER(Rule), c_uint32, c_uint64, c_int,
                                 POINTER(Summary)]
        L.oracle_tc_bfs.restype = c_int
        L.oracle_tc_bfs.argtypes = [u32, u32, c_uint64, c_uint32, c_int, POINTER(BfsResult), c_void_p, u32, u32,
                                    c_uint64]
        L.oracle_general_tc.restype = c_uint64
        L.oracle_general_tc.argtypes = [u32, u32, c_uint64, POINTER(c_uint64), POINTER(c_uint64), POINTER(c_uint64),
                                        u32, u32, u32, u32, c_uint64]
        L.oracle_load_binary_u32.restype = c_uint64
        L.oracle_load_binary_u32.argtypes = [ctypes.c_char_p, c_uint64, u32, u32, c_uint64, POINTER(LoadError)]
        _lib = L
    return _lib


class LoadError(ctypes.Structure):
    _fields_ = [("kind", c_int), ("line", c_uint64), ("column", c_uint32), ("value_offset", c_uint64),
                ("value_len", c_uint64)]


LOAD_ERRORS = {1: "GitLfsPointer", 2: "InvalidFieldCount", 3: "InvalidU32", 4: "Capacity"}


def load_binary_u32(text: bytes):
    """(first, second, None) or (None, None, error dict) for an in-memory
    text (storage/src/io.rs:72-141)."""
    cap = text.count(b"\n") + 1
    first = np.empty(cap, np.uint32)
    second = np.empty(cap, np.uint32)
    err = LoadError()
    n = lib().oracle_load_binary_u32(text, len(text), _p(first), _p(second), cap, ctypes.byref(err))
    if err.kind:
        e = {"kind": LOAD_ERRORS[err.kind], "line": err.line}
        if err.kind == 3:
            e["column"] = err.column
            e["value"] = text[err.value_offset:err.value_offset + err.value_len].decode("utf-8", "replace")
        return None, None, e
    return first[:n], second[:n], None


def _u32(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint32))


def _p(a: np.ndarray):
    return a.ctypes.data_as(c_void_p) if a.size else None


def distinct(first, second, threads: int = 1):
    a, b = _u32(first), _u32(second)
    oa, ob = np.empty(max(a.size, 1), np.uint32), np.empty(max(a.size, 1), np.uint32)
    n = lib().oracle_distinct(_p(a), _p(b), a.size, _p(oa), _p(ob), threads)
    return oa[:n], ob[:n]


def anti_join(left, right, threads: int = 1):
    lf, ls = _u32(left[0]), _u32(left[1])
    rf, rs = _u32(right[0]), _u32(right[1])
    oa, ob = np.empty(max(lf.size, 1), np.uint32), np.empty(max(lf.size, 1), np.uint32)
    n = lib().oracle_anti_join(_p(lf), _p(ls), lf.size, _p(rf), _p(rs), rf.size, _p(oa), _p(ob), threads)
    return oa[:n], ob[:n]


def union(left, right, threads: int = 1):
    lf, ls = _u32(left[0]), _u32(left[1])
    rf, rs = _u32(right[0]), _u32(right[1])
    t = max(lf.size + rf.size, 1)
    oa, ob = np.empty(t, np.uint32), np.empty(t, np.uint32)
    n = lib().oracle_union(_p(lf), _p(ls), lf.size, _p(rf), _p(rs), rf.size, _p(oa), _p(ob), threads)
    return oa[:n], ob[:n]


def range_index(values):
    v = _u32(values)
    keys = np.empty(max(v.size, 1), np.uint32)
    starts = np.empty(v.size + 1, np.uint32)
    rows = np.empty(max(v.size, 1), np.uint32)
    k = lib().oracle_range_index(_p(v), v.size, _p(keys), starts.ctypes.data_as(c_void_p), _p(rows))
    return keys[:k], starts[: k + 1], rows[: v.size]


def join(left, right, left_key: int, right_key: int, output, threads: int = 1, index=None):
    """output = ((side, column), (side, column)); side 0 left, 1 right.
    `index` = range_index(right[right_key]) when already built (the
    reference's crossover harness builds it outside the timed join,
    join-crossover.rs:127-133)."""
    left = (_u32(left[0]), _u32(left[1]))
    right = (_u32(right[0]), _u32(right[1]))
    keys, starts, rows = index if index is not None else range_index(right[right_key])
    lk = left[left_key]
    L = lib()
    m = L.oracle_join_count(_p(lk), lk.size, _p(keys), starts.ctypes.data_as(c_void_p), keys.size, threads)
    o0, o1 = np.empty(max(m, 1), np.uint32), np.empty(max(m, 1), np.uint32)
    p = [(left if side == 0 else right)[col] for side, col in output]
    L.oracle_join(_p(lk), lk.size, _p(keys), starts.ctypes.data_as(c_void_p), _p(rows), keys.size, _p(p[0]),
                  output[0][0], _p(p[1]), output[1][0], _p(o0), _p(o1), threads)
    return o0[:m], o1[:m]


def filter_rows(values, op: int, operand: int, threads: int = 1):
    v = _u32(values)
    out = np.empty(max(v.size, 1), np.uint32)
    n = lib().oracle_filter(_p(v), v.size, op, operand, _p(out), threads)
    return out[:n]


def tc_bfs(src, dst, n_nodes: int, threads: int = 8, closure: bool = False, per_source: bool = False):
    s, d = _u32(src), _u32(dst)
    res = BfsResult()
    ps = np.zeros(n_nodes, np.uint64) if per_source else None
    cap = 0
    cf = cs = None
    if closure:
        probe = BfsResult()
        lib().oracle_tc_bfs(_p(s), _p(d), s.size, n_nodes, threads, ctypes.byref(probe), None, None, None, 0)
        cap = probe.pairs
        cf, cs = np.empty(max(cap, 1), np.uint32), np.empty(max(cap, 1), np.uint32)
    rc = lib().oracle_tc_bfs(_p(s), _p(d), s.size, n_nodes, threads, ctypes.byref(res),
                             ps.ctypes.data_as(c_void_p) if ps is not None else None,
                             _p(cf) if cf is not None else None, _p(cs) if cs is not None else None, cap)
    if rc != 0:
        raise ValueError(f"oracle_tc_bfs failed with {rc}")
    out = {"pairs": res.pairs, "digest": res.digest, "max_dist": res.max_dist}
    if closure:
        out["closure"] = (cf[:cap], cs[:cap])
    if per_source:
        out["per_source"] = ps
    return out


def general_tc(src, dst):
    """The general CPU path on path/edge/nonlocal; returns a dict."""
    s, d = _u32(src), _u32(dst)
    cap = int(s.size) * int(s.size) + 16
    cap = min(cap, 1 << 24)
    pf, ps = np.empty(cap, np.uint32), np.empty(cap, np.uint32)
    nf, ns = np.empty(cap, np.uint32), np.empty(cap, np.uint32)
    nl, it, nr = c_uint64(), c_uint64(), c_uint64()
    n = lib().oracle_general_tc(_p(s), _p(d), s.size, ctypes.byref(nl), ctypes.byref(it), ctypes.byref(nr),
                                _p(pf), _p(ps), _p(nf), _p(ns), cap)
    return {"path": (pf[:n], ps[:n]), "nonlocal": (nf[: nl.value], ns[: nl.value]), "iterations": it.value,
            "new_rows": nr.value}


class Store:
    """Oracle RelationStore + RecursiveExecutor."""

    def __init__(self):
        self._h = c_void_p(lib().oracle_store_new())

    def __del__(self):
        if self._h:
            lib().oracle_store_free(self._h)
            self._h = None

    def insert_static(self, id: int, rel) -> int:
        a, b = _u32(rel[0]), _u32(rel[1])
        return lib().oracle_store_insert_static(self._h, id, _p(a), _p(b), a.size)

    def insert_recursive(self, id: int, full, delta) -> int:
        a, b = _u32(full[0]), _u32(full[1])
        c, d = _u32(delta[0]), _u32(delta[1])
        return lib().oracle_store_insert_recursive(self._h, id, _p(a), _p(b), a.size, _p(c), _p(d), c.size)

    def get(self, id: int, which: int = 0):
        n = lib().oracle_store_get(self._h, id, which, None, None)
        a, b = np.empty(max(n, 1), np.uint32), np.empty(max(n, 1), np.uint32)
        lib().oracle_store_get(self._h, id, which, _p(a), _p(b))
        return a[:n], b[:n]

    def run(self, relations, rules, max_iterations: int, threads: int = 1):
        rel = (c_uint32 * max(1, len(relations)))(*relations)
        rr = (Rule * max(1, len(rules)))(*rules)
        summ = Summary()
        rc = lib().oracle_run(self._h, rel, len(relations), rr, len(rules), max_iterations, threads,
                              ctypes.byref(summ))
        return rc, summ


def rule_from_plan(rule) -> Rule:
    """Rule from a sparkalog_amd.plans.RecursiveRulePlan."""
    r = Rule()
    r.target, r.delta_input, r.right_input = int(rule.target), int(rule.delta_input), int(rule.right_input)
    r.left_key, r.right_key = rule.join.left_key, rule.join.right_key
    for i in range(2):
        r.out_input[i] = int(rule.join.output[i].input)
        r.out_column[i] = rule.join.output[i].column
        r.distinct_columns[i] = rule.distinct.columns[i]
        r.anti_left[i], r.anti_right[i] = rule.anti_join.left[i], rule.anti_join.right[i]
        r.union_left[i], r.union_right[i] = rule.union.left[i], rule.union.right[i]
    return r
