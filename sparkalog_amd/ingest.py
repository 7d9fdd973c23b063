"""Device text ingest: the reference's binary u32 loader
(crates/sparkalog-storage/src/io.rs:72-141, load_binary_u32) with the parse
on the GPU.

The file's bytes go to device memory once (through a pinned staging
buffer); the line count, the line starts and the two-pass parse run as HIP
kernels (sparkalog_hip_text_lines / sparkalog_hip_load_binary_u32), and the
columns stay on the device, ready for RelationStore.insert_* or the operator
API. Errors are the reference's LoadError variants (io.rs:7-20) with the
same precedence: a Git LFS pointer line anywhere, else the first bad line.
"""
from __future__ import annotations

import ctypes
import os
import re
from pathlib import Path

import numpy as np
import torch

from . import _native
from .execution import HipStream, Relation, _check, _ptr

LOAD_OK = (1 << 64) - 1
ERR_CAPACITY = 1
ERR_LFS_POINTER = 0
ERR_FIELD_COUNT = (1 << 62) | 1
ERR_U32_FIRST = (1 << 62) | 2
ERR_U32_SECOND = (1 << 62) | 3


class LoadError(Exception):
    """storage::io::LoadError (io.rs:7-20)."""


class GitLfsPointer(LoadError):
    def __init__(self, path):
        super().__init__(f"{path} is a Git LFS pointer, not materialized graph data")
        self.path = path


class InvalidFieldCount(LoadError):
    def __init__(self, line: int):
        super().__init__(f"line {line} does not contain exactly two fields")
        self.line = line


class InvalidU32(LoadError):
    def __init__(self, line: int, column: int, value: str):
        super().__init__(f"line {line}, column {column} is not a u32: {value!r}")
        self.line = line
        self.column = column
        self.value = value


def _field(text: np.ndarray, line: int, column: int) -> str:
    """Text of field `column` (1-based) of line `line` (1-based)."""
    nl = np.flatnonzero(text == 10)
    start = 0 if line == 1 else int(nl[line - 2]) + 1
    end = int(nl[line - 1]) if line - 1 < nl.size else text.size
    fields = [f for f in re.split(rb"[ \t\n\x0b\x0c\r]+", bytes(text[start:end])) if f]
    return fields[column - 1].decode("utf-8", "replace") if len(fields) >= column else ""


def _raise_for(code: int, text: np.ndarray, path) -> None:
    if code == LOAD_OK:
        return
    if code == ERR_CAPACITY:
        raise RuntimeError("device ingest: the text has more lines than were counted")
    kind = code & ((3 << 62) | 3)
    line = (code >> 2) & ((1 << 60) - 1)
    if kind == ERR_LFS_POINTER:
        raise GitLfsPointer(path)
    if kind == ERR_FIELD_COUNT:
        raise InvalidFieldCount(line)
    column = 1 if kind == ERR_U32_FIRST else 2
    raise InvalidU32(line, column, _field(text, line, column))


def parse_binary_u32(data, stream: HipStream | None = None, path="memory") -> Relation:
    """Parse an in-memory text (bytes, bytearray or a uint8 numpy array) on
    the device; returns its two u32 columns as device tensors."""
    stream = stream or HipStream()
    lib = _native.hip()
    host = np.frombuffer(data, dtype=np.uint8) if not isinstance(data, np.ndarray) else data
    n = int(host.size)
    dev = stream.device
    # a 16-byte aligned device copy of the bytes, through pinned memory
    text = torch.empty(max(n, 16), dtype=torch.uint8, device=dev)
    if n:
        staging = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        staging.numpy()[:] = host
        with torch.cuda.stream(stream.torch_stream):
            text[:n].copy_(staging, non_blocking=True)
    words = torch.zeros(3, dtype=torch.int64, device=dev)  # lines, rows, error
    tb = ctypes.c_size_t(0)
    _check("sparkalog_hip_text_lines_temporary_bytes",
           lib.sparkalog_hip_text_lines_temporary_bytes(n, ctypes.byref(tb)))
    temp = torch.empty(max(tb.value, 1), dtype=torch.uint8, device=dev)
    _check("sparkalog_hip_text_lines",
           lib.sparkalog_hip_text_lines(_ptr(text), n, _ptr(words), _ptr(temp), tb.value, stream.raw))
    stream.synchronize()
    lines = int(words[0].item())
    first = torch.empty(max(lines, 1), dtype=torch.int32, device=dev)
    second = torch.empty(max(lines, 1), dtype=torch.int32, device=dev)
    _check("sparkalog_hip_load_binary_u32_temporary_bytes",
           lib.sparkalog_hip_load_binary_u32_temporary_bytes(n, lines, ctypes.byref(tb)))
    temp = torch.empty(max(tb.value, 1), dtype=torch.uint8, device=dev)
    rows_p = ctypes.c_void_p(words.data_ptr() + 8)
    err_p = ctypes.c_void_p(words.data_ptr() + 16)
    _check("sparkalog_hip_load_binary_u32",
           lib.sparkalog_hip_load_binary_u32(_ptr(text), n, lines, _ptr(first), _ptr(second), rows_p, err_p,
                                             _ptr(temp), tb.value, stream.raw))
    stream.synchronize()
    w = words.cpu().numpy().view(np.uint64)
    _raise_for(int(w[2]), host, path)
    rows = int(w[1])
    return Relation((first[:rows], second[:rows]))


def load_binary_u32(path, stream: HipStream | None = None) -> Relation:
    """load_binary_u32 (io.rs:72-83): the file's two u32 columns on the
    device, in file order (neither sorted nor deduplicated, as in the
    reference)."""
    path = Path(path)
    size = os.path.getsize(path)
    data = np.fromfile(path, dtype=np.uint8, count=size) if size else np.empty(0, np.uint8)
    return parse_binary_u32(data, stream, path)
