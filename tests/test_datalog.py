"""The Datalog frontend in front of the device pipeline (sparkalog_amd.datalog):
parsing, interning, stratification and lowering on the CPU (the reference's
own frontend tests restated), and on the GPU the Database running programs
on the device, checked against the oracle and the BFS fixtures."""
import numpy as np
import pytest

import oracle
from sparkalog_amd import datalog as dl
from sparkalog_amd import workload


def _resolve(source):
    prog, diags = dl.parse_program(source)
    assert diags == []
    cat = dl.ProgramCatalog()
    resolved, diags = dl.resolve_program(prog, cat)
    assert diags == []
    return resolved, cat


# ---- parser (parser.rs:512-572) ----

def test_parses_facts_rules_negation_constants_and_outputs():
    prog, diags = dl.parse_program(
        "edge('a, \"b\").\nweight('a, 42).\npath(x, y) :- edge(x, y), !blocked(y).\n.output path\n")
    assert diags == []
    assert len(prog.rules) == 3
    assert prog.rules[0].head.terms[0].value == dl.Symbol("a")
    assert isinstance(prog.rules[0].head.terms[1].value, dl.Text)
    assert prog.rules[1].head.terms[1].value == 42
    assert prog.rules[2].body[1].negated
    assert prog.outputs[0][0] == "path"


def test_recovers_at_statement_boundaries():
    prog, diags = dl.parse_program("good(1, 2). bad(1 2). also(3, 4).")
    assert len(diags) == 1
    assert [r.head.predicate for r in prog.rules] == ["good", "also"]


def test_reports_lexical_errors_without_losing_valid_rules():
    prog, diags = dl.parse_program("ok(1, 2). n(99999999999, 1). s(\"open, 1). tail(5, 6).")
    assert any("does not fit in u32" in d.message for d in diags)
    assert prog.rules[0].head.predicate == "ok"


def test_accepts_souffle_declarations_and_comments():
    prog, diags = dl.parse_program("""
        // comment
        .decl edge(from:symbol, to:symbol)
        # another
        .decl path(from:symbol, to:symbol)
        % and another
        .input edge
        .output path
        path(x, y) :- edge(x, y).
    """)
    assert diags == []
    assert [(n, a) for n, a, _ in prog.declarations] == [("edge", 2), ("path", 2)]
    assert prog.inputs[0][0] == "edge" and prog.outputs[0][0] == "path"


def test_string_escapes():
    prog, diags = dl.parse_program('s("a\\"b\\\\c\\n", 1).')
    assert diags == []
    assert str(prog.rules[0].head.terms[0].value) == 'a"b\\c\n'


# ---- resolution (resolve.rs) ----

def test_values_intern_in_first_seen_order():
    # subset.rs:474-497: edges 1->2->3->4 become ids 0..3
    resolved, cat = _resolve("edge(1, 2). edge(2, 3). edge(3, 4).")
    assert [[t.value for t in r.head.terms] for r in resolved.rules] == [[0, 1], [1, 2], [2, 3]]
    assert [cat.values.get(i) for i in range(4)] == [1, 2, 3, 4]


def test_symbol_string_and_number_are_distinct_values():
    _, cat = _resolve("v('a, \"a\"). w(1, 'a).")
    assert len(cat.values) == 3
    assert cat.values.get(0) == "a" and isinstance(cat.values.get(0), dl.Symbol)
    assert isinstance(cat.values.get(1), dl.Text)
    assert cat.values.get(2) == 1


def test_unsafe_rules_are_diagnosed():
    cat = dl.ProgramCatalog()
    prog, _ = dl.parse_program("bad(x) :- !missing(x).")
    _, diags = dl.resolve_program(prog, cat)
    assert diags and "not bound by a positive body atom" in diags[0].message
    prog, _ = dl.parse_program("p(x, y) :- e(x, z).")
    _, diags = dl.resolve_program(prog, dl.ProgramCatalog())
    assert "head variable y" in diags[0].message


def test_arity_conflicts_are_diagnosed():
    prog, _ = dl.parse_program("e(1, 2). e(1, 2, 3).")
    _, diags = dl.resolve_program(prog, dl.ProgramCatalog())
    assert "arity" in diags[0].message


def test_vectorised_interning_matches_row_by_row():
    rng = np.random.default_rng(4)
    a = rng.integers(0, 50, 500, dtype=np.uint32)
    b = rng.integers(0, 50, 500, dtype=np.uint32)
    one, many = dl.ValueCatalog(), dl.ValueCatalog()
    one.intern(7)
    many.intern(7)
    ids = [(one.intern(int(x)), one.intern(int(y))) for x, y in zip(a, b)]
    fa, fb = many.intern_u32_rows(a, b)
    assert list(zip(fa.tolist(), fb.tolist())) == ids
    assert [one.get(i) for i in range(len(one))] == [many.get(i) for i in range(len(many))]


# ---- stratification (schedule.rs) ----

def test_negation_goes_to_a_later_stratum():
    resolved, cat = _resolve("""
        edge(1, 2).
        path(x, y) :- edge(x, y).
        path(x, z) :- path(x, y), edge(y, z).
        nonlocal(x, z) :- path(x, z), !edge(x, z).
    """)
    strata = dl.stratify(resolved)
    name = lambda p: cat.predicates.get(p).name  # noqa: E731
    flat = [(i, [name(p) for p in scc.predicates], scc.recursive) for i, s in enumerate(strata) for scc in s]
    assert ([n for _, n, _ in flat].index(["path"]) < [n for _, n, _ in flat].index(["nonlocal"]))
    assert next(r for _, n, r in flat if n == ["path"])
    stratum_of = {tuple(n): i for i, n, _ in flat}
    assert stratum_of[("nonlocal",)] > stratum_of[("path",)]


def test_negative_cycle_is_rejected():
    resolved, _ = _resolve("e(1, 2). p(x, y) :- e(x, y), !q(x, y). q(x, y) :- e(x, y), !p(x, y).")
    with pytest.raises(dl.StratificationError):
        dl.stratify(resolved)


def test_mutual_recursion_is_one_component():
    resolved, cat = _resolve("e(1, 2). a(x, y) :- e(x, y). a(x, z) :- b(x, y), e(y, z). "
                             "b(x, z) :- a(x, y), e(y, z).")
    sccs = [scc for s in dl.stratify(resolved) for scc in s if scc.recursive]
    assert len(sccs) == 1
    assert {cat.predicates.get(p).name for p in sccs[0].predicates} == {"a", "b"}


# ---- lowering (subset.rs:39-97, extended) ----

def test_tc_lowers_to_a_seed_and_a_recursive_rule():
    resolved, cat = _resolve("edge(1, 2). edge(2, 3). edge(3, 4). path(x, y) :- edge(x, y). "
                             "path(x, z) :- path(x, y), edge(y, z). .output path")
    plan = dl.lower_binary_hip(resolved)
    step = plan.steps[0]
    assert isinstance(step, dl.RecursiveStep)
    assert len(step.seeds) == 1 and isinstance(step.seeds[0][1], dl.ProjectRule)
    rule = step.plan.rules[0]
    # the reference's TC plan (generic.rs:629-661): left_key 1, right_key 0, out [L.0, R.1]
    assert (rule.join.left_key, rule.join.right_key) == (1, 0)
    assert [(int(p.input), p.column) for p in rule.join.output] == [(0, 0), (1, 1)]


def test_nonrecursive_rules_lower_to_device_operators():
    # the reference rejects this program (subset.rs:56-63, test :499-507)
    resolved, _ = _resolve("""
        edge(1, 2).
        copy(x, y) :- edge(x, y).
        flip(y, x) :- edge(x, y).
        two(x, z) :- edge(x, y), edge(y, z).
        path(x, y) :- edge(x, y).
        path(x, z) :- path(x, y), edge(y, z).
        nonlocal(x, z) :- path(x, z), !edge(x, z).
        rev_miss(x, z) :- path(x, z), !flip(z, x).
    """)
    plan = dl.lower_binary_hip(resolved)
    kinds = [type(s.rules[0]).__name__ for s in plan.steps if isinstance(s, dl.DerivedStep)]
    assert sorted(kinds) == ["JoinRule", "NegationRule", "NegationRule", "ProjectRule", "ProjectRule"]
    neg = [s.rules[0] for s in plan.steps if isinstance(s, dl.DerivedStep) and
           isinstance(s.rules[0], dl.NegationRule)]
    assert {r.key_columns for r in neg} == {(0, 1), (1, 0)}


@pytest.mark.parametrize("source,message", [
    ("t(1, 2, 3). p(x, y) :- t(x, y, z).", "binary"),
    ("e(1, 2). p(x, y) :- e(x, 5), e(y, 5).", "constants"),
    ("e(1, 2). p(x, x) :- e(x, x).", "repeated"),
    ("e(1, 2). p(x, y) :- e(x, y). p(x, z) :- p(x, y), p(y, z).", "one recursive body atom"),
    ("e(1, 2). p(x, z) :- e(x, y), e(y, z), e(z, x).", "bodies of one atom"),
    ("e(1, 2). p(x, y) :- e(x, y), e(y, x).", "exactly one variable"),
])
def test_unsupported_shapes_are_explicit(source, message):
    resolved, _ = _resolve(source)
    with pytest.raises(dl.LoweringError, match=message):
        dl.lower_binary_hip(resolved)


def test_database_frontend_errors():
    db = dl.Database()
    with pytest.raises(dl.ParseFailed):
        db.load_program("bad(x :- nope.")
    with pytest.raises(dl.ResolveFailed):
        db.load_program("bad(x) :- !missing(x).")
    with pytest.raises(dl.LoweringFailed):
        db.load_program("t(1, 2, 3). pair(x, z) :- t(x, y, z). .output pair")
    with pytest.raises(dl.NoProgram):
        db.run()


def test_database_insert_checks():
    db = dl.Database()
    db.load_program("edge('a, 'b). path(x, y) :- edge(x, y). .output path")
    with pytest.raises(dl.UnknownPredicate):
        db.insert("nope", ["a", "b"])
    with pytest.raises(dl.ArityMismatch):
        db.insert("edge", ["a"])
    with pytest.raises(dl.NotRun):
        db.query("path")


def test_delimited_parser():
    rows = dl.parse_delimited('1,2\n"x,y",b\n\n3 , c\n')
    assert rows == [[1, 2], [dl.Text("x,y"), dl.Symbol("b")], [3, dl.Symbol("c")]]
    with pytest.raises(dl.DelimitedError, match="empty field"):
        dl.parse_delimited("1,\n")
    with pytest.raises(dl.DelimitedError, match="unterminated"):
        dl.parse_delimited('"open,1\n')


# ---- on the GPU ----

@pytest.mark.gpu
def test_database_decodes_outputs(stream):
    # database.rs:391-410
    db = dl.Database(stream)
    db.load_program("edge('a, 'b). path(x, y) :- edge(x, y). .output path")
    summary = db.run()
    assert summary.backend == dl.ExecutionBackend.BINARY_HIP
    out = db.outputs()
    assert [o.predicate for o in out] == ["path"]
    assert out[0].rows == [[dl.Symbol("a"), dl.Symbol("b")]]
    assert all(isinstance(v, dl.Symbol) for v in out[0].rows[0])


@pytest.mark.gpu
def test_database_inserted_facts_extend_the_next_run(stream):
    # database.rs:412-426
    db = dl.Database(stream)
    db.load_program("edge('a, 'b). path(x, y) :- edge(x, y). path(x, z) :- path(x, y), edge(y, z). .output path")
    db.run()
    db.insert("edge", ["b", "c"])
    with pytest.raises(dl.NotRun):
        db.query("path")
    db.run()
    assert db.query("path").rows == [[dl.Symbol("a"), dl.Symbol("b")], [dl.Symbol("a"), dl.Symbol("c")],
                                     [dl.Symbol("b"), dl.Symbol("c")]]


@pytest.mark.gpu
def test_interned_tc_matches_reference_vector(stream):
    # subset.rs:474-497: 6 rows over ids 0..3, 3 iterations, decoded to 1..4
    db = dl.Database(stream)
    db.load_program("edge(1, 2). edge(2, 3). edge(3, 4). path(x, y) :- edge(x, y). "
                    "path(x, z) :- path(x, y), edge(y, z). .output path")
    summary = db.run()
    assert summary.sccs[0].iterations == 3
    rel = db.execution.relations[db.catalog.predicates.id("path")]
    first, second = rel.to_host()
    assert first.tolist() == [0, 0, 0, 1, 1, 2] and second.tolist() == [1, 2, 3, 2, 3, 3]
    assert db.query("path").rows == [[1, 2], [1, 3], [1, 4], [2, 3], [2, 4], [3, 4]]


@pytest.mark.gpu
def test_full_program_with_negation_matches_general_path(expected, stream):
    # C1's full program: TC then nonlocal = path \ edge, every stratum on the
    # device; checked against the general path's restatement and the fixture
    e = expected["random_dag_1000_4000"]
    src, dst = workload.random_dag()
    db = dl.Database(stream)
    facts = "".join(f"edge({a}, {b}).\n" for a, b in zip(src.tolist(), dst.tolist()))
    db.load_program(facts + """
        path(x, y) :- edge(x, y).
        path(x, z) :- path(x, y), edge(y, z).
        nonlocal(x, z) :- path(x, z), !edge(x, z).
        .output path
        .output nonlocal
    """)
    summary = db.run()
    assert summary.sccs[0].iterations == e["max_dist"]
    path, nonlocal_ = db.outputs()
    assert len(path.rows) == e["path_rows"] and len(nonlocal_.rows) == e["nonlocal_rows"]
    got_path = np.array(path.rows, np.uint32)
    got_nl = np.array(nonlocal_.rows, np.uint32)
    assert workload.pair_digest(got_path[:, 0], got_path[:, 1]) == e["path_digest"]
    assert workload.pair_digest(got_nl[:, 0], got_nl[:, 1]) == e["nonlocal_digest"]
    # the general path's semi-naive restatement agrees row for row (as sets)
    general = oracle.general_tc(src, dst)
    assert general["iterations"] == summary.sccs[0].iterations
    assert {tuple(r) for r in path.rows} == set(zip(*(c.tolist() for c in general["path"])))
    assert {tuple(r) for r in nonlocal_.rows} == set(zip(*(c.tolist() for c in general["nonlocal"])))


@pytest.mark.gpu
def test_nonrecursive_rules_on_the_device(stream):
    import itertools

    rng = np.random.default_rng(11)
    edges = sorted({(int(a), int(b)) for a, b in rng.integers(0, 40, (300, 2))})
    db = dl.Database(stream)
    facts = "".join(f"edge({a}, {b}).\n" for a, b in edges)
    db.load_program(facts + """
        flip(y, x) :- edge(x, y).
        two(x, z) :- edge(x, y), edge(y, z).
        both(x, y) :- edge(x, y).
        both(x, y) :- flip(x, y).
        oneway(x, y) :- edge(x, y), !flip(x, y).
        .output two
    """)
    db.run()
    E = set(edges)
    F = {(b, a) for a, b in E}
    two = {(a, d) for (a, b), (c, d) in itertools.product(E, E) if b == c}
    dec = lambda name: {tuple(r) for r in db.query(name).rows}  # noqa: E731
    assert dec("flip") == F
    assert dec("two") == two
    assert dec("both") == E | F
    assert dec("oneway") == E - F
    # rows come back sorted by interned id pairs, as FULL is stored
    ids = db.execution.relations[db.catalog.predicates.id("two")].to_host()
    packed = (ids[0].astype(np.uint64) << np.uint64(32)) | ids[1]
    assert np.all(np.diff(packed.astype(np.int64)) > 0)


@pytest.mark.gpu
def test_database_loads_files_on_the_device(tmp_path, stream):
    path = tmp_path / "edge.txt"
    path.write_text("# chain\n10 20\n20 30\n\n30 40\n")
    db = dl.Database(stream)
    db.load_program(".decl edge(a:number, b:number) .input edge "
                    "path(x, y) :- edge(x, y). path(x, z) :- path(x, y), edge(y, z). .output path")
    assert db.load_binary_u32("edge", path) == 3
    assert db.load_delimited("edge", "40,50\n") == 1
    db.run()
    assert db.query("path").rows == [[10, 20], [10, 30], [10, 40], [10, 50], [20, 30], [20, 40], [20, 50],
                                     [30, 40], [30, 50], [40, 50]]


# ---- random programs against a naive stratified evaluator ----

def _naive_eval(facts, rules, order):
    """Set-semantics evaluation of binary rules stratum by stratum (order =
    lists of predicates, each list a stratum evaluated to its fixpoint).
    Rules: (head, body) with body a list of (negated, pred, (t0, t1)), terms
    variable names. Independent of the device pipeline and of the oracle."""
    rel = {p: set(v) for p, v in facts.items()}
    for stratum in order:
        changed = True
        while changed:
            changed = False
            for head, hv, body in rules:
                if head not in stratum:
                    continue
                envs = [{}]
                for neg, pred, tv in body:
                    if neg:
                        continue
                    nxt = []
                    for env in envs:
                        for a, b in rel.get(pred, ()):
                            e = dict(env)
                            ok = True
                            for var, val in zip(tv, (a, b)):
                                if e.setdefault(var, val) != val:
                                    ok = False
                            if ok:
                                nxt.append(e)
                    envs = nxt
                for neg, pred, tv in body:
                    if neg:
                        envs = [e for e in envs if (e[tv[0]], e[tv[1]]) not in rel.get(pred, set())]
                for e in envs:
                    t = (e[hv[0]], e[hv[1]])
                    if t not in rel.setdefault(head, set()):
                        rel[head].add(t)
                        changed = True
    return rel


def _random_program(rng):
    """A random stratified program over edb e0, e1 and idb p0..p3 with rule
    shapes the device pipeline lowers."""
    facts = {f"e{i}": {(int(a), int(b)) for a, b in rng.integers(0, 12, (int(rng.integers(5, 40)), 2))}
             for i in range(2)}
    rules = []
    defined = ["e0", "e1"]
    strata = []
    for k in range(4):
        p = f"p{k}"
        shape = rng.choice(["copy", "flip", "join", "neg", "tc", "tc_left"])
        a, b = rng.choice(defined, 2)
        if shape == "copy":
            rules.append((p, ("x", "y"), [(False, a, ("x", "y"))]))
        elif shape == "flip":
            rules.append((p, ("y", "x"), [(False, a, ("x", "y"))]))
        elif shape == "join":
            rules.append((p, ("x", "z"), [(False, a, ("x", "y")), (False, b, ("y", "z"))]))
        elif shape == "neg":
            rules.append((p, ("x", "y"), [(False, a, ("x", "y")), (True, b, ("y", "x") if rng.random() < 0.5
                                                                              else ("x", "y"))]))
        elif shape == "tc":
            rules.append((p, ("x", "y"), [(False, a, ("x", "y"))]))
            rules.append((p, ("x", "z"), [(False, p, ("x", "y")), (False, b, ("y", "z"))]))
        else:  # recursion through the left column of the static input
            rules.append((p, ("x", "y"), [(False, a, ("x", "y"))]))
            rules.append((p, ("x", "z"), [(False, b, ("x", "y")), (False, p, ("y", "z"))]))
        if rng.random() < 0.3:  # a second rule for the same predicate
            rules.append((p, ("x", "y"), [(False, b, ("y", "x"))]))
        defined.append(p)
        strata.append([p])
    return facts, rules, strata


def _source(facts, rules):
    lines = [f"{p}({a}, {b})." for p, ts in facts.items() for a, b in sorted(ts)]
    for head, hv, body in rules:
        lits = ", ".join(("!" if neg else "") + f"{pred}({t[0]}, {t[1]})" for neg, pred, t in body)
        lines.append(f"{head}({hv[0]}, {hv[1]}) :- {lits}.")
    return "\n".join(lines) + "\n"


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(24))
def test_random_programs_match_naive_evaluation(seed, stream):
    rng = np.random.default_rng(0xDA7A + seed)
    facts, rules, strata = _random_program(rng)
    expected = _naive_eval(facts, rules, strata)
    db = dl.Database(stream)
    db.load_program(_source(facts, rules))
    db.run()
    for p in ["e0", "e1", "p0", "p1", "p2", "p3"]:
        got = {tuple(r) for r in db.query(p).rows} if db.catalog.predicates.id(p) in db.active else set()
        assert got == expected.get(p, set()), (seed, p, _source(facts, rules))
