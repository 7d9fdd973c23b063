"""The autotuner's search and tuning files (CPU), and on the GPU: results do
not depend on the tuning file, the segmented distinct's tile, or the other
knobs the tuner sets; invalid option values are refused."""
import json

import numpy as np
import pytest

from sparkalog_amd import autotune


def _bowl(optimum, signature=("same",)):
    """A separable cost with its minimum at `optimum` (knob -> value)."""
    calls = []

    def measure(values):
        calls.append(dict(values))
        cost = 100.0
        for k, v in values.items():
            cands = autotune.KNOBS[k]
            cost += 10.0 * abs(cands.index(v) - cands.index(optimum.get(k, cands[0])))
        return cost, signature

    return measure, calls


def test_descent_finds_a_separable_minimum():
    optimum = {"distinct_tile": 4096, "merge_ratio": 6, "indexed_anti_join": 1}
    measure, calls = _bowl(optimum)
    best, ms, trace = autotune.coordinate_descent(measure)
    for k, v in optimum.items():
        assert best[k] == v
    assert ms == 100.0
    # each configuration measured once
    keys = [tuple(sorted(c.items())) for c in calls]
    assert len(keys) == len(set(keys)) == len(trace)


def test_descent_keeps_defaults_without_a_real_gain():
    measure, _ = _bowl({})
    best, _, _ = autotune.coordinate_descent(measure)
    assert best == autotune.DEFAULTS


def test_descent_skips_graph_max_rows_without_graph_mode():
    measure, calls = _bowl({})
    autotune.coordinate_descent(measure)
    assert all(c["graph_max_rows"] == autotune.DEFAULTS["graph_max_rows"] for c in calls if not c["graph"])


def test_descent_refuses_a_knob_that_changes_results():
    def measure(values):
        return 1.0, ("tile", values["distinct_tile"])

    with pytest.raises(AssertionError, match="changed the result"):
        autotune.coordinate_descent(measure)


def test_tuning_file_round_trip(tmp_path, monkeypatch):
    data = {"format": "sparkalog-hip-tuning/1", "device": "x", "arch": "gfx950",
            "profiles": {"closure": {"values": {"distinct_tile": 2048, "graph": 1}},
                         "batched": {"values": {"merge_ratio": 4}}}}
    path = tmp_path / "t.json"
    autotune.write(data, path)
    assert autotune.read(path) == data
    assert autotune.profile_values(path, "closure") == {"distinct_tile": 2048, "graph": 1}
    assert autotune.profile_values(str(path), "batched") == {"merge_ratio": 4}
    assert autotune.profile_values(path, "other") == {}
    assert autotune.profile_values(None, "closure") == {}
    assert autotune.profile_values({"merge_ratio": 2}, "closure") == {"merge_ratio": 2}
    monkeypatch.setenv(autotune.ENV_FILE, str(path))
    assert autotune.default_file() == path
    monkeypatch.setenv(autotune.ENV_FILE, "")
    assert autotune.default_file() is None


def test_tuning_file_rejects_unknown_knobs(tmp_path):
    path = tmp_path / "bad.json"
    path.write_text(json.dumps({"format": "sparkalog-hip-tuning/1",
                                "profiles": {"closure": {"values": {"warp_size": 32}}}}))
    with pytest.raises(ValueError, match="unknown knob"):
        autotune.read(path)
    path.write_text(json.dumps({"profiles": {}}))
    with pytest.raises(ValueError, match="not a tuning file"):
        autotune.read(path)


def test_apply_sets_each_knob():
    class Fake:
        def __init__(self):
            self.set = {}

        def set_option(self, name, value):
            self.set[name] = value

    f = Fake()
    autotune.apply(f, {"distinct_tile": 1024, "graph": 1})
    assert f.set == {"distinct_tile": 1024, "graph": 1}
    with pytest.raises(KeyError):
        autotune.apply(f, {"nonsense": 1})


def test_package_file_is_valid_when_present():
    if not autotune.PACKAGE_FILE.exists():
        pytest.skip("no tuning file in the package")
    data = autotune.read(autotune.PACKAGE_FILE)
    assert set(data["profiles"]) <= set(autotune.PROFILES)
    for prof in data["profiles"].values():
        for k, v in prof["values"].items():
            assert v in autotune.KNOBS[k]


# ---- on the GPU ----

def _closure_arrays(store):
    from sparkalog_amd import plans, workload
    from sparkalog_amd.recursion import RecursiveExecutor

    R = plans.RelationId
    s, d, _ = workload.component_graph(nodes=1 << 14)
    store.insert_static(R(0), (s, d))
    store.insert_recursive(R(1), (s, d), (s, d))
    summ = RecursiveExecutor.compile(plans.transitive_closure_scc(R(1), R(0)), store).run(store, 10_000)
    return store.full(R(1)), summ.iterations, (s, d)


@pytest.mark.gpu
def test_results_do_not_depend_on_the_tuning_file(tmp_path, monkeypatch, expected, stream):
    import oracle
    from sparkalog_amd.recursion import RelationStore

    e = expected["component_graph_16384"]
    files = []
    for i, values in enumerate([{"distinct_tile": 1024, "merge_ratio": 2, "indexed_anti_join": 1,
                                 "sparse_probe_rows": 1 << 18},
                                {"distinct_tile": 4096, "merge_ratio": 8, "graph": 1, "graph_max_rows": 1 << 22},
                                {"segmented_distinct": 0, "sparse_probe_rows": 1 << 24}]):
        path = tmp_path / f"t{i}.json"
        autotune.write({"format": "sparkalog-hip-tuning/1", "profiles": {"closure": {"values": values}}}, path)
        files.append(path)
    ref = None
    for path in [None] + files:
        monkeypatch.setenv(autotune.ENV_FILE, str(path) if path else "")
        store = RelationStore(stream)
        assert store.tuning == (autotune.profile_values(path, "closure") if path else {})
        (a, b), iterations, (s, d) = _closure_arrays(store)
        if ref is None:
            bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
            assert np.array_equal(a, bfs["closure"][0]) and np.array_equal(b, bfs["closure"][1])
            assert iterations == e["max_dist"]
            ref = (a, b)
        assert np.array_equal(a, ref[0]) and np.array_equal(b, ref[1]), path
        assert iterations == e["max_dist"]
        store.close()


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [512, 1024, 2048, 4096])
def test_distinct_tile_does_not_change_results(tile, stream):
    import oracle
    from sparkalog_amd.recursion import RelationStore

    store = RelationStore(stream, tuning=None)
    store.set_option("distinct_tile", tile)
    (a, b), _, (s, d) = _closure_arrays(store)
    bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
    assert np.array_equal(a, bfs["closure"][0]) and np.array_equal(b, bfs["closure"][1])


@pytest.mark.gpu
def test_batched_profile_applies_to_batched_waves(tmp_path, stream):
    import oracle
    from sparkalog_amd import batch, workload
    from sparkalog_amd.recursion import RelationStore

    path = tmp_path / "b.json"
    autotune.write({"format": "sparkalog-hip-tuning/1",
                    "profiles": {"batched": {"values": {"distinct_tile": 2048, "merge_ratio": 4}}}}, path)
    store = RelationStore(stream, tuning=path, profile="batched")
    assert store.tuning == {"distinct_tile": 2048, "merge_ratio": 4}
    s, d, offs, counts = workload.batch_graphs(seed=0x7A, instances=96, n_max=1024)
    res = batch.run_batched(s, d, offs, counts, store, budget_pairs=1e12)
    for i in range(counts.size):
        r = oracle.tc_bfs(s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]], int(counts[i]))
        assert (int(res.rows[i]), int(res.digests[i])) == (r["pairs"], r["digest"]), i


@pytest.mark.gpu
def test_invalid_option_values_are_refused(stream):
    from sparkalog_amd.recursion import GenericExecutionError, RelationStore

    store = RelationStore(stream, tuning=None)
    for name, value in [("distinct_tile", 256), ("distinct_tile", 8192), ("distinct_tile", 1000),
                        ("merge_ratio", 1), ("segmented_distinct", 2), ("indexed_anti_join", 3),
                        ("graph", 2), ("graph_max_rows", 1 << 40), ("join_index", 5), ("no_such_knob", 1)]:
        with pytest.raises(GenericExecutionError, match="InvalidArgument"):
            store.set_option(name, value)


@pytest.mark.gpu
@pytest.mark.parametrize("profile", autotune.PROFILES)
def test_package_tuning_matches_bfs(profile, expected, stream):
    # the committed tuning file, as bench.py and a default store apply it
    import oracle
    from sparkalog_amd import batch, workload
    from sparkalog_amd.recursion import RelationStore

    if not autotune.PACKAGE_FILE.exists():
        pytest.skip("no tuning file in the package")
    store = RelationStore(stream, tuning=autotune.PACKAGE_FILE, profile=profile)
    if profile == "closure":
        (a, b), iterations, (s, d) = _closure_arrays(store)
        bfs = oracle.tc_bfs(s, d, 1 << 14, closure=True)
        assert np.array_equal(a, bfs["closure"][0]) and np.array_equal(b, bfs["closure"][1])
        assert iterations == expected["component_graph_16384"]["max_dist"]
    else:
        s, d, offs, counts = workload.batch_graphs(seed=0x7B, instances=96, n_max=1024)
        res = batch.run_batched(s, d, offs, counts, store, budget_pairs=1e12)
        for i in range(counts.size):
            r = oracle.tc_bfs(s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]], int(counts[i]))
            assert (int(res.rows[i]), int(res.digests[i])) == (r["pairs"], r["digest"]), i
