"""Device text ingest (sparkalog_amd.ingest) against the oracle's restatement
of load_binary_u32 (crates/sparkalog-storage/src/io.rs:72-141).

The oracle is pinned on the CPU by the reference's own tests (io.rs:148-174),
by texts whose values the generator wrote (ground truth independent of any
parser), and by known answers of Rust's str::parse::<u32> and str::trim /
split_whitespace. The GPU tests compare the device parse with the oracle,
exactly, on the same texts and on every error kind, plus large generated
texts against their ground truth."""
import numpy as np
import pytest

import oracle
from sparkalog_amd import workload

# io.rs:148-174
REFERENCE_CASES = [
    (b"# source destination\n1\t2\n\n3 4\n% comment\n5\t6\n", ([1, 3, 5], [2, 4, 6]), None),
    (b"version https://git-lfs.github.com/spec/v1\noid sha256:abc\nsize 12\n", None,
     {"kind": "GitLfsPointer", "line": 1}),
    (b"1 2 3\n", None, {"kind": "InvalidFieldCount", "line": 1}),
]

# known answers of Rust's parse/trim/split rules
RULE_CASES = [
    (b"", ([], []), None),
    (b"\n\n \t\n", ([], []), None),
    (b"1 2", ([1], [2]), None),                                   # no final newline
    (b"1 2\r\n3 4\r\n", ([1, 3], [2, 4]), None),                  # CRLF
    (b"+5 007\n", ([5], [7]), None),                              # '+' sign, leading zeros
    (b"4294967295 0\n", ([4294967295], [0]), None),               # u32::MAX
    (b"\x0b1\x0c \t2 \r\n", ([1], [2]), None),                     # VT and FF are whitespace
    (b"  # 1 2\n% 3 4\n7 8\n", ([7], [8]), None),                 # comments after leading spaces
    (b"4294967296 1\n", None, {"kind": "InvalidU32", "line": 1, "column": 1, "value": "4294967296"}),
    (b"1 -2\n", None, {"kind": "InvalidU32", "line": 1, "column": 2, "value": "-2"}),
    (b"1 +\n", None, {"kind": "InvalidU32", "line": 1, "column": 2, "value": "+"}),
    (b"1 ++2\n", None, {"kind": "InvalidU32", "line": 1, "column": 2, "value": "++2"}),
    (b"0x10 1\n", None, {"kind": "InvalidU32", "line": 1, "column": 1, "value": "0x10"}),
    (b"1_000 1\n", None, {"kind": "InvalidU32", "line": 1, "column": 1, "value": "1_000"}),
    (b"# c\n\n1 x\n", None, {"kind": "InvalidU32", "line": 3, "column": 2, "value": "x"}),
    (b"1\n", None, {"kind": "InvalidFieldCount", "line": 1}),
    (b"1\x1c2\n", None, {"kind": "InvalidFieldCount", "line": 1}),  # \x1c is not whitespace in Rust
    (b"x y z\n", None, {"kind": "InvalidFieldCount", "line": 1}),   # field count before parse
    (b"a 1\nb c\n", None, {"kind": "InvalidU32", "line": 1, "column": 1, "value": "a"}),
    # a pointer line anywhere wins over an earlier bad line (io.rs:75-79)
    (b"1 2 3\nversion https://git-lfs.github.com/spec/v1\n", None, {"kind": "GitLfsPointer", "line": 2}),
    # ... but only as a whole trimmed line
    (b"version https://git-lfs.github.com/spec/v1 x\n", None, {"kind": "InvalidFieldCount", "line": 1}),
]


def _check_oracle(text, rows, err):
    f, s, e = oracle.load_binary_u32(text)
    if err is None:
        assert e is None, e
        assert f.tolist() == rows[0] and s.tolist() == rows[1]
    else:
        assert e == err


@pytest.mark.parametrize("text,rows,err", REFERENCE_CASES + RULE_CASES)
def test_oracle_known_answers(text, rows, err):
    _check_oracle(text, rows, err)


@pytest.mark.parametrize("style", [0, 1, 2, 4, 8, 16, 31])
def test_oracle_matches_generated_ground_truth(style):
    text, a, b = workload.edge_text(0x1E + style, 5000, key_space=0, style=style)
    f, s, e = oracle.load_binary_u32(bytes(text))
    assert e is None
    assert np.array_equal(f, a) and np.array_equal(s, b)


def test_generator_agrees_with_python_parsing_on_plain_text():
    # with plain separators, Python's own split/int reads the same values
    text, a, b = workload.edge_text(0x77, 2000, key_space=1 << 20, style=0)
    rows = [tuple(int(x) for x in line.split()) for line in bytes(text).decode().splitlines()]
    assert rows == list(zip(a.tolist(), b.tolist()))


def test_generator_is_deterministic():
    t1, a1, _ = workload.edge_text(9, 100, style=31)
    t2, a2, _ = workload.edge_text(9, 100, style=31)
    assert np.array_equal(t1, t2) and np.array_equal(a1, a2)
    assert t1[-1] != 10  # style 16: no final newline


# ---- on the GPU ----

def _device(text, stream):
    from sparkalog_amd import ingest

    return ingest.parse_binary_u32(np.frombuffer(bytes(text), np.uint8) if not isinstance(text, np.ndarray)
                                   else text, stream)


def _host(rel):
    return (rel.columns[0].cpu().numpy().view(np.uint32), rel.columns[1].cpu().numpy().view(np.uint32))


@pytest.mark.gpu
@pytest.mark.parametrize("case", range(len(REFERENCE_CASES + RULE_CASES)))
def test_device_known_answers(case, stream):
    from sparkalog_amd import ingest

    text, rows, err = (REFERENCE_CASES + RULE_CASES)[case]
    if err is None:
        f, s = _host(_device(text, stream))
        assert f.tolist() == rows[0] and s.tolist() == rows[1]
        return
    expected = {"GitLfsPointer": ingest.GitLfsPointer, "InvalidFieldCount": ingest.InvalidFieldCount,
                "InvalidU32": ingest.InvalidU32}[err["kind"]]
    with pytest.raises(expected) as info:
        _device(text, stream)
    if err["kind"] != "GitLfsPointer":
        assert info.value.line == err["line"]
    if err["kind"] == "InvalidU32":
        assert (info.value.column, info.value.value) == (err["column"], err["value"])


@pytest.mark.gpu
@pytest.mark.parametrize("rows,style", [(1, 0), (1, 16), (63, 31), (4095, 1), (4096, 2), (4097, 31),
                                        (100_000, 31), (1 << 20, 15)])
def test_device_matches_oracle(rows, style, stream):
    # sizes around the 4,096-line parse tile and many 16 KiB text tiles
    text, a, b = workload.edge_text(0x100 + rows + style, rows, key_space=0, style=style)
    f, s = _host(_device(text, stream))
    of, os_, e = oracle.load_binary_u32(bytes(text))
    assert e is None
    assert np.array_equal(f, of) and np.array_equal(s, os_)
    assert np.array_equal(f, a) and np.array_equal(s, b)


@pytest.mark.gpu
def test_device_long_lines_and_tile_edges(stream):
    # lines longer than a 16 KiB text tile, lines ending exactly at tile
    # boundaries, and runs of blank lines across tiles
    parts = [b" " * 20000 + b"1 2\n", b"3" + b" " * 40000 + b"4\n", b"\n" * 30000, b"#" * 16383 + b"\n"]
    filler = b"5 6\n"
    body = b"".join(parts)
    pad = (16384 - len(body) % 16384) % 16384
    text = body + b"7" + b" " * max(pad - 4, 1) + b"8\n" + filler * 5000
    f, s = _host(_device(text, stream))
    of, os_, e = oracle.load_binary_u32(text)
    assert e is None and np.array_equal(f, of) and np.array_equal(s, os_)
    assert f[:3].tolist() == [1, 3, 7] and s[:3].tolist() == [2, 4, 8]


@pytest.mark.gpu
@pytest.mark.parametrize("where", [0, 4095, 4096, 99_999])
def test_device_reports_the_first_bad_line(where, stream):
    from sparkalog_amd import ingest

    text, a, b = workload.edge_text(0xBAD, 100_000, key_space=1 << 16, style=0)
    lines = bytes(text).split(b"\n")
    lines[where] = b"12 x4"
    if where + 7 < 100_000:
        lines[where + 7] = b"1 2 3"  # a later error does not win
    bad = b"\n".join(lines)
    _, _, e = oracle.load_binary_u32(bad)
    assert e == {"kind": "InvalidU32", "line": where + 1, "column": 2, "value": "x4"}
    with pytest.raises(ingest.InvalidU32) as info:
        _device(bad, stream)
    assert (info.value.line, info.value.column, info.value.value) == (where + 1, 2, "x4")


@pytest.mark.gpu
def test_device_large_text_ground_truth(stream):
    # 2^24 rows (about 200 MB of text with every style): rows and an
    # order-dependent digest against the generator's values
    rows = 1 << 24
    text, a, b = workload.edge_text(0x1A6E, rows, key_space=0, style=31)
    f, s = _host(_device(text, stream))
    assert f.size == rows
    assert np.array_equal(f, a) and np.array_equal(s, b)


@pytest.mark.gpu
def test_loaded_edges_feed_the_executor(tmp_path, expected, stream):
    # file -> device columns -> sorted unique (device distinct) -> closure
    from sparkalog_amd import execution, ingest, plans
    from sparkalog_amd.recursion import RecursiveExecutor, RelationStore

    src, dst = workload.random_dag()
    perm = np.random.default_rng(3).permutation(src.size)
    path = tmp_path / "edge.facts"
    lines = ["# random DAG"] + [f"{x}\t{y}" for x, y in zip(src[perm], dst[perm])]
    lines += [f"{src[0]} {dst[0]}"]  # a duplicate
    path.write_text("\n".join(lines) + "\n")
    rel = ingest.load_binary_u32(path, stream)
    assert len(rel) == src.size + 1
    edges = execution.distinct_hip(rel, plans.BinaryDistinct((0, 1)), stream)
    store = RelationStore(stream)
    store.insert_static(plans.RelationId(0), edges.columns)
    store.insert_recursive(plans.RelationId(1), edges.columns, edges.columns)
    RecursiveExecutor.compile(plans.transitive_closure_scc(plans.RelationId(1), plans.RelationId(0)),
                              store).run(store, 10_000)
    digest, sorted_unique, rows = store.digest(plans.RelationId(1))
    e = expected["random_dag_1000_4000"]
    assert (rows, digest, sorted_unique) == (e["path_rows"], e["path_digest"], True)
