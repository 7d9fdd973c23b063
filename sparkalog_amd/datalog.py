"""Datalog programs on the device pipeline: the reference's frontend
(crates/sparkalog-datalog) in front of the HIP operators and executor.

  parse_program     parser.rs:42-510 (facts, rules, negation, constants,
                    .decl / .input / .output, comments)
  resolve_program   resolve.rs:21-250 (predicate and value interning in
                    first-seen order, rule safety)
  stratify          schedule.rs:45-322 (SCCs of the dependency graph,
                    negation only across strata)
  lower_binary_hip  subset.rs:39-97 extended: the reference lowers a program
                    to the binary pipeline only when every derived predicate
                    is recursive (rejection at :56-63); here non-recursive
                    binary rules lower too, each to device operators -
                    projection (distinct), a two-atom join (join + distinct)
                    and stratified negation (anti-join against a relation of
                    an earlier stratum) - so a program such as
                        path(x, y) :- edge(x, y).
                        path(x, z) :- path(x, y), edge(y, z).
                        nonlocal(x, z) :- path(x, z), !edge(x, z).
                    runs entirely on the GPU: TC in the semi-naive executor,
                    then nonlocal as a SortedBinaryAntiJoin.
  execute_binary_hip  subset.rs:369-414: facts deduplicated and sorted on
                    the device (the reference's BTreeSet), each stratum in
                    order, relations kept in HBM between strata
  Database          database.rs:19-290 with ExecutionBackend BINARY_HIP:
                    load_program / insert / load_delimited / load_binary_u32
                    (device text ingest) / run / query / outputs, query rows
                    read back through pinned staging and decoded from
                    interned ids.

There is no CPU evaluator behind this frontend: a program outside the device
subset (arity other than 2, constants or repeated variables in rule bodies,
bodies other than one atom, two joined atoms or one atom and one negated
atom, recursive rules the executor cannot take) is refused at load with a
LoweringError naming the rule.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field
from enum import Enum

import numpy as np

from .plans import (BinaryDistinct, BinaryEqualityJoin, JoinInput, JoinProjection, RecursiveRulePlan,
                    RecursiveSccPlan, RelationId, SortedBinaryAntiJoin, SortedBinaryUnion)

U32_MAX = 0xFFFFFFFF


# ---------------------------------------------------------------- values

class Symbol(str):
    """InternedValue::Symbol ('name in the source)."""

    def __repr__(self) -> str:
        return f"'{str(self)}"


class Text(str):
    """InternedValue::String ("..." in the source)."""

    def __repr__(self) -> str:
        return f'Text({str(self)!r})'


def _value_key(value) -> tuple:
    if isinstance(value, Text):
        return ("string", str(value))
    if isinstance(value, str):  # Symbol, or a plain str (From<&str> is a Symbol, catalog.rs:24-34)
        return ("symbol", str(value))
    if isinstance(value, (int, np.integer)) and not isinstance(value, bool) and 0 <= int(value) <= U32_MAX:
        return ("u32", int(value))
    raise TypeError(f"{value!r} is not a u32, a symbol or a string")


def _from_key(key: tuple):
    kind, v = key
    return v if kind == "u32" else (Symbol(v) if kind == "symbol" else Text(v))


class CatalogError(Exception):
    pass


class ValueCatalog:
    """catalog.rs:95-140: values interned to dense u32 ids in first-seen order."""

    def __init__(self):
        self._ids: dict[tuple, int] = {}
        self._values: list[tuple] = []

    def __len__(self) -> int:
        return len(self._values)

    def intern(self, value) -> int:
        key = _value_key(value)
        vid = self._ids.get(key)
        if vid is None:
            if len(self._values) > U32_MAX:
                raise CatalogError("value catalog exceeds u32 ids")
            vid = len(self._values)
            self._ids[key] = vid
            self._values.append(key)
        return vid

    def get(self, vid: int):
        return _from_key(self._values[vid])

    def intern_u32_rows(self, first: np.ndarray, second: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        """Intern two u32 columns row by row (first then second value of each
        row, as inserting the rows one at a time does), vectorised."""
        flat = np.stack([np.asarray(first, np.uint32), np.asarray(second, np.uint32)], axis=1).ravel()
        if flat.size == 0:
            return np.empty(0, np.uint32), np.empty(0, np.uint32)
        uniq, first_at, inverse = np.unique(flat, return_index=True, return_inverse=True)
        ids = np.empty(uniq.size, np.int64)
        for j in np.argsort(first_at, kind="stable"):
            ids[j] = self.intern(int(uniq[j]))
        mapped = ids[inverse].astype(np.uint32)
        return mapped[0::2].copy(), mapped[1::2].copy()


@dataclass
class PredicateMetadata:
    name: str
    arity: int | None = None


class PredicateCatalog:
    """catalog.rs:42-93."""

    def __init__(self):
        self._by_name: dict[str, int] = {}
        self._predicates: list[PredicateMetadata] = []

    def intern(self, name: str) -> int:
        pid = self._by_name.get(name)
        if pid is None:
            pid = len(self._predicates)
            self._by_name[name] = pid
            self._predicates.append(PredicateMetadata(name))
        return pid

    def declare(self, name: str, arity: int) -> int:
        pid = self.intern(name)
        meta = self._predicates[pid]
        if meta.arity is None:
            meta.arity = arity
        elif meta.arity != arity:
            raise CatalogError(f"predicate {name} has arity {meta.arity}, not {arity}")
        return pid

    def id(self, name: str) -> int | None:
        return self._by_name.get(name)

    def get(self, pid: int) -> PredicateMetadata:
        return self._predicates[pid]


@dataclass
class ProgramCatalog:
    predicates: PredicateCatalog = field(default_factory=PredicateCatalog)
    values: ValueCatalog = field(default_factory=ValueCatalog)


# ---------------------------------------------------------------- syntax

@dataclass(frozen=True)
class Span:
    start: int = 0
    end: int = 0

    def join(self, other: "Span") -> "Span":
        return Span(min(self.start, other.start), max(self.end, other.end))


@dataclass(frozen=True)
class Diagnostic:
    message: str
    span: Span

    def line_column(self, source: str) -> tuple[int, int]:
        before = source[:self.span.start]
        return before.count("\n") + 1, self.span.start - (before.rfind("\n") + 1) + 1


@dataclass(frozen=True)
class Token:
    kind: str      # ident symbol string number ( ) , . ! :- : invalid
    value: object
    span: Span


_IDENT = re.compile(r"[A-Za-z_][A-Za-z0-9_]*")
_SYMBOL = re.compile(r"[A-Za-z0-9_]+")
_ESCAPES = {"n": "\n", "r": "\r", "t": "\t", '"': '"', "\\": "\\"}


def lex(source: str) -> tuple[list[Token], list[Diagnostic]]:
    """parser.rs:79-258."""
    tokens: list[Token] = []
    diags: list[Diagnostic] = []
    i, n = 0, len(source)
    while i < n:
        c = source[i]
        if c in " \t\n\r\x0b\x0c":
            i += 1
        elif c in "#%" or source.startswith("//", i):
            while i < n and source[i] != "\n":
                i += 1
        elif c in "(),.!":
            tokens.append(Token(c, None, Span(i, i + 1)))
            i += 1
        elif c == ":":
            if source.startswith(":-", i):
                tokens.append(Token(":-", None, Span(i, i + 2)))
                i += 2
            else:
                tokens.append(Token(":", None, Span(i, i + 1)))
                i += 1
        elif c == "'":
            m = _SYMBOL.match(source, i + 1)
            if not m:
                diags.append(Diagnostic("expected a symbol name after apostrophe", Span(i, i + 1)))
                i += 1
            else:
                tokens.append(Token("symbol", m.group(), Span(i, m.end())))
                i = m.end()
        elif c == '"':
            start, i, value, closed = i, i + 1, [], False
            while i < n:
                ch = source[i]
                if ch == '"':
                    i += 1
                    closed = True
                    break
                if ch == "\\":
                    if i + 1 >= n:
                        i += 1
                        break
                    esc = source[i + 1]
                    if esc not in _ESCAPES:
                        diags.append(Diagnostic("unsupported string escape", Span(i, i + 2)))
                    value.append(_ESCAPES.get(esc, esc))
                    i += 2
                else:
                    value.append(ch)
                    i += 1
            if closed:
                tokens.append(Token("string", "".join(value), Span(start, i)))
            else:
                diags.append(Diagnostic("unterminated string literal", Span(start, i)))
        elif c.isascii() and c.isdigit():
            start = i
            while i < n and source[i].isascii() and source[i].isdigit():
                i += 1
            v = int(source[start:i])
            if v > U32_MAX:
                diags.append(Diagnostic("integer literal does not fit in u32", Span(start, i)))
                tokens.append(Token("invalid", None, Span(start, i)))
            else:
                tokens.append(Token("number", v, Span(start, i)))
        else:
            m = _IDENT.match(source, i)
            if m:
                tokens.append(Token("ident", m.group(), Span(i, m.end())))
                i = m.end()
            else:
                diags.append(Diagnostic(f"unexpected character {c!r}", Span(i, i + 1)))
                tokens.append(Token("invalid", None, Span(i, i + 1)))
                i += 1
    return tokens, diags


@dataclass(frozen=True)
class SourceTerm:
    variable: str | None = None  # a variable name, or None for a constant
    value: object = None         # int, Symbol or Text
    span: Span = Span()


@dataclass(frozen=True)
class SourceAtom:
    predicate: str
    terms: tuple[SourceTerm, ...]
    span: Span


@dataclass(frozen=True)
class SourceLiteral:
    negated: bool
    atom: SourceAtom
    span: Span


@dataclass(frozen=True)
class SourceRule:
    head: SourceAtom
    body: tuple[SourceLiteral, ...]
    span: Span


@dataclass
class SourceProgram:
    rules: list[SourceRule] = field(default_factory=list)
    declarations: list[tuple[str, int, Span]] = field(default_factory=list)
    inputs: list[tuple[str, Span]] = field(default_factory=list)
    outputs: list[tuple[str, Span]] = field(default_factory=list)


class _ParseError(Exception):
    def __init__(self, diagnostic: Diagnostic):
        self.diagnostic = diagnostic


class _Parser:
    """parser.rs:259-510 (statement-level recovery: skip past the next '.')."""

    def __init__(self, tokens: list[Token], end: int):
        self.t = tokens
        self.i = 0
        self.end = end

    def peek(self) -> Token | None:
        return self.t[self.i] if self.i < len(self.t) else None

    def advance(self) -> Token | None:
        tok = self.peek()
        if tok is not None:
            self.i += 1
        return tok

    def check(self, kind: str) -> bool:
        tok = self.peek()
        return tok is not None and tok.kind == kind

    def error_here(self, message: str) -> _ParseError:
        tok = self.peek()
        return _ParseError(Diagnostic(message, tok.span if tok else Span(self.end, self.end)))

    def expect(self, kind: str, message: str) -> Token:
        if not self.check(kind):
            raise self.error_here(message)
        return self.advance()

    def program(self) -> tuple[SourceProgram, list[Diagnostic]]:
        prog, diags = SourceProgram(), []
        while self.peek() is not None:
            try:
                if self.check("."):
                    self.directive(prog)
                else:
                    prog.rules.append(self.rule())
            except _ParseError as e:
                diags.append(e.diagnostic)
                self.synchronize()
        return prog, diags

    def synchronize(self) -> None:
        while self.peek() is not None:
            if self.advance().kind == ".":
                return

    def rule(self) -> SourceRule:
        head = self.atom()
        body = []
        if self.check(":-"):
            self.advance()
            while True:
                body.append(self.literal())
                if not self.check(","):
                    break
                self.advance()
        dot = self.expect(".", "expected `.` after rule")
        return SourceRule(head, tuple(body), head.span.join(dot.span))

    def literal(self) -> SourceLiteral:
        bang = self.advance() if self.check("!") else None
        atom = self.atom()
        return SourceLiteral(bang is not None, atom, bang.span.join(atom.span) if bang else atom.span)

    def atom(self) -> SourceAtom:
        name = self.expect("ident", "expected a predicate name")
        self.expect("(", "expected `(` after predicate name")
        terms = []
        if not self.check(")"):
            while True:
                terms.append(self.term())
                if not self.check(","):
                    break
                self.advance()
        right = self.expect(")", "expected `)` after atom")
        return SourceAtom(name.value, tuple(terms), name.span.join(right.span))

    def term(self) -> SourceTerm:
        tok = self.advance()
        if tok is None:
            raise self.error_here("expected a variable or constant")
        if tok.kind == "ident":
            return SourceTerm(variable=tok.value, span=tok.span)
        if tok.kind == "symbol":
            return SourceTerm(value=Symbol(tok.value), span=tok.span)
        if tok.kind == "string":
            return SourceTerm(value=Text(tok.value), span=tok.span)
        if tok.kind == "number":
            return SourceTerm(value=tok.value, span=tok.span)
        raise _ParseError(Diagnostic("expected a variable or constant", tok.span))

    def directive(self, prog: SourceProgram) -> None:
        dot = self.expect(".", "expected directive")
        name = self.expect("ident", "expected directive name after `.`")
        if name.value in ("output", "input"):
            pred = self.expect("ident", f"expected predicate name after `.{name.value}`")
            (prog.outputs if name.value == "output" else prog.inputs).append((pred.value, pred.span))
        elif name.value == "decl":
            pred = self.expect("ident", "expected predicate name after `.decl`")
            self.expect("(", "expected `(` in `.decl`")
            arity = 0
            if not self.check(")"):
                while True:
                    self.expect("ident", "expected attribute name in `.decl`")
                    self.expect(":", "expected `:` after attribute name")
                    self.expect("ident", "expected attribute type in `.decl`")
                    arity += 1
                    if not self.check(","):
                        break
                    self.advance()
            right = self.expect(")", "expected `)` after `.decl`")
            prog.declarations.append((pred.value, arity, dot.span.join(right.span)))
        else:
            raise _ParseError(Diagnostic(f"unknown directive .{name.value}", name.span))


def parse_program(source: str) -> tuple[SourceProgram, list[Diagnostic]]:
    tokens, diags = lex(source)
    prog, more = _Parser(tokens, len(source)).program()
    return prog, diags + more


# ---------------------------------------------------------------- resolution

@dataclass(frozen=True)
class Term:
    variable: int | None = None  # variable id within its rule
    value: int | None = None     # interned value id


@dataclass(frozen=True)
class Atom:
    predicate: int
    terms: tuple[Term, ...]
    span: Span


@dataclass(frozen=True)
class Literal:
    negated: bool
    atom: Atom


@dataclass(frozen=True)
class Rule:
    head: Atom
    body: tuple[Literal, ...]
    variable_names: tuple[str, ...]
    span: Span

    def is_fact(self) -> bool:
        return not self.body


@dataclass
class ResolvedProgram:
    declarations: list[int] = field(default_factory=list)
    inputs: list[int] = field(default_factory=list)
    rules: list[Rule] = field(default_factory=list)
    outputs: list[int] = field(default_factory=list)


def resolve_program(source: SourceProgram, catalog: ProgramCatalog) -> tuple[ResolvedProgram, list[Diagnostic]]:
    """resolve.rs:21-104 and :106-250."""
    diags: list[Diagnostic] = []
    prog = ResolvedProgram()
    declared = set()
    for name, arity, span in source.declarations:
        if name in declared:
            diags.append(Diagnostic(f"predicate {name} is declared more than once", span))
            continue
        declared.add(name)
        try:
            prog.declarations.append(catalog.predicates.declare(name, arity))
        except CatalogError as e:
            diags.append(Diagnostic(str(e), span))
    seen_inputs = set()
    for name, span in source.inputs:
        if name in seen_inputs:
            diags.append(Diagnostic(f"input predicate {name} is listed more than once", span))
            continue
        seen_inputs.add(name)
        pid = catalog.predicates.id(name)
        if pid is not None and pid in prog.declarations:
            prog.inputs.append(pid)
        else:
            diags.append(Diagnostic(f"input predicate {name} has no declaration", span))
    for rule in source.rules:
        resolved, errors = _resolve_rule(rule, catalog)
        if errors:
            diags.extend(errors)
        else:
            prog.rules.append(resolved)
    for rule in prog.rules:
        if not rule.is_fact() and rule.head.predicate in prog.inputs:
            diags.append(Diagnostic("an input (EDB) predicate cannot also be derived by a rule", rule.head.span))
    for name, span in source.outputs:
        pid = catalog.predicates.id(name)
        if pid is not None and catalog.predicates.get(pid).arity is not None:
            prog.outputs.append(pid)
        else:
            diags.append(Diagnostic(f"output predicate {name} is not defined", span))
    return prog, diags


def _resolve_rule(rule: SourceRule, catalog: ProgramCatalog) -> tuple[Rule | None, list[Diagnostic]]:
    names: list[str] = []
    ids: dict[str, int] = {}
    diags: list[Diagnostic] = []

    def atom(src: SourceAtom) -> Atom | None:
        try:
            pid = catalog.predicates.declare(src.predicate, len(src.terms))
        except CatalogError as e:
            diags.append(Diagnostic(str(e), src.span))
            pid = None
        terms = []
        for t in src.terms:
            if t.variable is not None:
                if t.variable not in ids:
                    ids[t.variable] = len(names)
                    names.append(t.variable)
                terms.append(Term(variable=ids[t.variable]))
            else:
                try:
                    terms.append(Term(value=catalog.values.intern(t.value)))
                except CatalogError as e:
                    diags.append(Diagnostic(str(e), t.span))
        return None if pid is None else Atom(pid, tuple(terms), src.span)

    head = atom(rule.head)
    body = []
    for lit in rule.body:
        a = atom(lit.atom)
        if a is not None:
            body.append(Literal(lit.negated, a))
    positive = {t.variable for lit in body if not lit.negated for t in lit.atom.terms if t.variable is not None}
    if head is not None:
        for t, st in zip(head.terms, rule.head.terms):
            if t.variable is not None and t.variable not in positive:
                diags.append(Diagnostic(f"head variable {names[t.variable]} is not bound by a positive body atom",
                                        st.span))
    for lit, slit in zip(body, rule.body):
        if not lit.negated:
            continue
        for t, st in zip(lit.atom.terms, slit.atom.terms):
            if t.variable is not None and t.variable not in positive:
                diags.append(Diagnostic(f"variable {names[t.variable]} in a negated atom is not bound by a "
                                        "positive body atom", st.span))
    if diags:
        return None, diags
    return Rule(head, tuple(body), tuple(names), rule.span), []


# ---------------------------------------------------------------- stratification

class StratificationError(Exception):
    pass


@dataclass(frozen=True)
class ScheduledScc:
    predicates: tuple[int, ...]
    recursive: bool


def _predicate_count(prog: ResolvedProgram) -> int:
    ids = [r.head.predicate for r in prog.rules] + [l.atom.predicate for r in prog.rules for l in r.body]
    ids += prog.declarations + prog.outputs
    return max(ids) + 1 if ids else 0


def stratify(prog: ResolvedProgram) -> list[list[ScheduledScc]]:
    """schedule.rs:45-145: strata in order, each a list of SCCs in
    dependency order; a negative dependency inside an SCC is an error."""
    deps = []
    for r in prog.rules:
        for lit in r.body:
            d = (r.head.predicate, lit.atom.predicate, lit.negated)
            if d not in deps:
                deps.append(d)
    n = _predicate_count(prog)
    active = [False] * n
    for r in prog.rules:
        active[r.head.predicate] = True
        for lit in r.body:
            active[lit.atom.predicate] = True
    for p in prog.outputs + prog.declarations:
        active[p] = True
    graph = [[] for _ in range(n)]
    reverse = [[] for _ in range(n)]
    for head, body, _ in deps:
        graph[head].append(body)
        reverse[body].append(head)
    # Kosaraju: finishing order over the graph, then components on the reverse
    order, seen = [], [False] * n
    for start in range(n):
        if not active[start] or seen[start]:
            continue
        stack = [(start, iter(graph[start]))]
        seen[start] = True
        while stack:
            node, it = stack[-1]
            nxt = next(it, None)
            if nxt is None:
                order.append(node)
                stack.pop()
            elif not seen[nxt]:
                seen[nxt] = True
                stack.append((nxt, iter(graph[nxt])))
    comp_of = [-1] * n
    comps: list[list[int]] = []
    for node in reversed(order):
        if comp_of[node] >= 0:
            continue
        cid = len(comps)
        comps.append([])
        todo = [node]
        comp_of[node] = cid
        while todo:
            v = todo.pop()
            comps[cid].append(v)
            for w in reverse[v]:
                if comp_of[w] < 0:
                    comp_of[w] = cid
                    todo.append(w)
    for head, body, negated in deps:
        if negated and comp_of[head] == comp_of[body]:
            raise StratificationError(f"predicate {head} depends negatively on {body} within one recursive "
                                      "component")
    strata_of = [0] * len(comps)
    for _ in range(len(comps)):
        changed = False
        for head, body, negated in deps:
            h, b = comp_of[head], comp_of[body]
            if h != b and strata_of[h] < strata_of[b] + int(negated):
                strata_of[h] = strata_of[b] + int(negated)
                changed = True
        if not changed:
            break
    # components in dependency order: a component after every one it reads
    indeg = {c: set() for c in range(len(comps))}
    for head, body, _ in deps:
        if comp_of[head] != comp_of[body]:
            indeg[comp_of[head]].add(comp_of[body])
    topo, done = [], set()
    while len(topo) < len(comps):
        ready = sorted((c for c in range(len(comps)) if c not in done and indeg[c] <= done),
                       key=lambda c: min(comps[c]))
        topo.append(ready[0])
        done.add(ready[0])
    strata: list[list[ScheduledScc]] = [[] for _ in range(max(strata_of, default=-1) + 1)]
    for c in topo:
        preds = tuple(sorted(comps[c]))
        recursive = len(preds) > 1 or (preds[0], preds[0], False) in deps
        strata[strata_of[c]].append(ScheduledScc(preds, recursive))
    return [s for s in strata if s]


# ---------------------------------------------------------------- lowering

class LoweringError(Exception):
    def __init__(self, message: str, span: Span = Span()):
        super().__init__(message)
        self.message = message
        self.span = span


@dataclass(frozen=True)
class ProjectRule:
    """head(h0, h1) :- source(...): the source's columns `columns`."""

    source: int
    columns: tuple[int, int]


@dataclass(frozen=True)
class JoinRule:
    """head :- left(..), right(..) sharing one variable."""

    left: int
    right: int
    join: BinaryEqualityJoin


@dataclass(frozen=True)
class NegationRule:
    """head :- positive(..), !negative(..): the positive relation's columns
    in the negated atom's order (`key_columns`) are anti-joined against the
    negative relation; the head takes the survivors' columns `head_columns`
    (indexes into the keyed pair)."""

    positive: int
    key_columns: tuple[int, int]
    negative: int
    head_columns: tuple[int, int]


DerivedRule = ProjectRule | JoinRule | NegationRule


@dataclass(frozen=True)
class DerivedStep:
    """A non-recursive predicate: the union of its rules' results and facts."""

    target: int
    rules: tuple[DerivedRule, ...]


@dataclass(frozen=True)
class RecursiveStep:
    """A recursive SCC (subset.rs:117-157): seeds from its non-recursive
    rules, then the semi-naive executor over its recursive rules."""

    plan: RecursiveSccPlan
    seeds: tuple[tuple[int, DerivedRule], ...]


@dataclass
class HipProgramPlan:
    predicate_count: int
    facts: dict[int, list[tuple[int, int]]]
    steps: list[DerivedStep | RecursiveStep]
    outputs: list[int]


def _vars_of(atom: Atom, rule: Rule, what: str) -> list[int]:
    out = []
    for t in atom.terms:
        if t.variable is None:
            raise LoweringError(f"{what}: constants in rule bodies are not lowered to the device pipeline",
                                atom.span)
        out.append(t.variable)
    if len(set(out)) != len(out):
        raise LoweringError(f"{what}: a variable repeated within one atom is not lowered to the device pipeline",
                            atom.span)
    return out


def _head_vars(rule: Rule) -> list[int]:
    out = []
    for t in rule.head.terms:
        if t.variable is None:
            raise LoweringError("rule heads must project variables", rule.head.span)
        out.append(t.variable)
    return out


def _single_join_key(left: list[int], right: list[int], span: Span) -> tuple[int, int]:
    keys = [(i, j) for i, a in enumerate(left) for j, b in enumerate(right) if a == b]
    if len(keys) != 1:
        raise LoweringError("joined atoms must share exactly one variable", span)
    return keys[0]


def _projection(head: list[int], left: list[int], right: list[int], span: Span):
    out = []
    for v in head:
        if v in left:
            out.append(JoinProjection(JoinInput.LEFT, left.index(v)))
        elif v in right:
            out.append(JoinProjection(JoinInput.RIGHT, right.index(v)))
        else:
            raise LoweringError("head variable is absent from the rule body", span)
    return tuple(out)


def _lower_derived_rule(rule: Rule) -> DerivedRule:
    head = _head_vars(rule)
    pos = [l for l in rule.body if not l.negated]
    neg = [l for l in rule.body if l.negated]
    if len(pos) == 1 and not neg:
        src = _vars_of(pos[0].atom, rule, "projection")
        return ProjectRule(pos[0].atom.predicate, (src.index(head[0]), src.index(head[1])))
    if len(pos) == 2 and not neg:
        lv = _vars_of(pos[0].atom, rule, "join")
        rv = _vars_of(pos[1].atom, rule, "join")
        lk, rk = _single_join_key(lv, rv, rule.span)
        return JoinRule(pos[0].atom.predicate, pos[1].atom.predicate,
                        BinaryEqualityJoin(lk, rk, _projection(head, lv, rv, rule.head.span)))
    if len(pos) == 1 and len(neg) == 1:
        pv = _vars_of(pos[0].atom, rule, "negation")
        nv = _vars_of(neg[0].atom, rule, "negation")
        key = (pv.index(nv[0]), pv.index(nv[1]))  # bound by the positive atom (resolve.rs:146-162)
        keyed = [pv[key[0]], pv[key[1]]]
        return NegationRule(pos[0].atom.predicate, key, neg[0].atom.predicate,
                            (keyed.index(head[0]), keyed.index(head[1])))
    raise LoweringError("the device pipeline lowers bodies of one atom, two joined atoms, or one atom and one "
                        "negated atom", rule.span)


def _lower_recursive_rule(rule: Rule, scc: set[int]) -> RecursiveRulePlan:
    """subset.rs:194-242."""
    if len(rule.body) != 2 or any(l.negated for l in rule.body):
        raise LoweringError("a recursive clause must contain two positive atoms", rule.span)
    rec = next(l for l in rule.body if l.atom.predicate in scc)
    right = next((l for l in rule.body if l.atom.predicate not in scc), None)
    if right is None:
        raise LoweringError("the initial subset requires a non-recursive right input", rule.span)
    lv = _vars_of(rec.atom, rule, "recursive join")
    rv = _vars_of(right.atom, rule, "recursive join")
    lk, rk = _single_join_key(lv, rv, rule.span)
    return RecursiveRulePlan(target=RelationId(rule.head.predicate), delta_input=RelationId(rec.atom.predicate),
                             right_input=RelationId(right.atom.predicate),
                             join=BinaryEqualityJoin(lk, rk, _projection(_head_vars(rule), lv, rv, rule.head.span)))


def lower_binary_hip(prog: ResolvedProgram) -> HipProgramPlan:
    """lower_binary (subset.rs:39-97) with non-recursive binary rules and
    negation strata lowered as well."""
    schedule = stratify(prog)
    for rule in prog.rules:  # validate_binary_arities, subset.rs:99-117
        if len(rule.head.terms) != 2:
            raise LoweringError("the device pipeline requires binary predicates", rule.head.span)
        for lit in rule.body:
            if len(lit.atom.terms) != 2:
                raise LoweringError("the device pipeline requires binary predicates", lit.atom.span)
    facts: dict[int, list[tuple[int, int]]] = {}
    for rule in prog.rules:
        if rule.is_fact():
            facts.setdefault(rule.head.predicate, []).append(tuple(t.value for t in rule.head.terms))
    steps: list[DerivedStep | RecursiveStep] = []
    for stratum in schedule:
        for scc in stratum:
            members = set(scc.predicates)
            rules = [r for r in prog.rules if r.head.predicate in members and not r.is_fact()]
            if not scc.recursive:
                if rules:
                    steps.append(DerivedStep(scc.predicates[0], tuple(_lower_derived_rule(r) for r in rules)))
                continue
            seeds, rec_rules = [], []
            for r in rules:
                count = sum(1 for l in r.body if l.atom.predicate in members)
                if count == 0:
                    seeds.append((r.head.predicate, _lower_derived_rule(r)))
                elif count == 1:
                    rec_rules.append(_lower_recursive_rule(r, members))
                else:
                    raise LoweringError("the initial executable subset permits one recursive body atom", r.span)
            plan = RecursiveSccPlan(tuple(RelationId(p) for p in scc.predicates), tuple(rec_rules))
            steps.append(RecursiveStep(plan, tuple(seeds)))
    return HipProgramPlan(_predicate_count(prog), facts, steps, list(prog.outputs))


def explain(plan: HipProgramPlan, catalog: ProgramCatalog) -> str:
    name = lambda p: catalog.predicates.get(int(p)).name  # noqa: E731
    lines = ["preferred backend: binary HIP (MI355X)"]
    for step in plan.steps:
        if isinstance(step, RecursiveStep):
            rels = ", ".join(name(r) for r in step.plan.relations)
            lines.append(f"recursive SCC [{rels}]: {len(step.seeds)} seed rule(s), {len(step.plan.rules)} "
                         "semi-naive rule(s) (join, distinct, anti-join, union per iteration)")
        else:
            kinds = ", ".join({ProjectRule: "projection", JoinRule: "join", NegationRule: "anti-join"}[type(r)]
                              for r in step.rules)
            lines.append(f"derived {name(step.target)}: {kinds}")
    return "\n".join(lines)


# ---------------------------------------------------------------- execution

@dataclass
class SccRunSummary:
    """GeneralSccSummary (general.rs) as database.rs:131-138 fills it."""

    relations: list[tuple[int, int]]
    iterations: int
    new_rows: int


@dataclass
class ProgramExecution:
    relations: dict
    summaries: list[SccRunSummary]


def execute_binary_hip(plan: HipProgramPlan, stream=None, max_iterations: int = 10_000,
                       extra_facts: dict | None = None) -> ProgramExecution:
    """execute_binary (subset.rs:369-414) on the device, stratum by stratum.
    `extra_facts` maps relation ids to (first, second) u32 arrays (interned
    ids) added to the program's facts."""
    import torch

    from . import execution as ex
    from .recursion import RecursiveExecutor, RelationStore

    stream = stream or ex.HipStream()
    dev = stream.device
    empty = lambda: ex.Relation((torch.empty(0, dtype=torch.int32, device=dev),) * 2)  # noqa: E731

    def upload(first, second) -> "ex.Relation":
        a = torch.from_numpy(np.ascontiguousarray(first, np.uint32).view(np.int32)).to(dev)
        b = torch.from_numpy(np.ascontiguousarray(second, np.uint32).view(np.int32)).to(dev)
        return ex.Relation((a, b))

    def distinct(rel):
        return rel if len(rel) == 0 else ex.distinct_hip(rel, BinaryDistinct(), stream)

    def union(a, b):
        if len(a) == 0:
            return b
        if len(b) == 0:
            return a
        return ex.union_hip(a, b, SortedBinaryUnion(), stream)

    rels: dict[int, ex.Relation] = {}
    for rid in set(plan.facts) | set(extra_facts or {}):
        parts = []
        if rid in plan.facts:
            arr = np.asarray(plan.facts[rid], np.uint32).reshape(-1, 2)
            parts.append((arr[:, 0], arr[:, 1]))
        if extra_facts and rid in extra_facts:
            parts.append(extra_facts[rid])
        first = np.concatenate([p[0] for p in parts])
        second = np.concatenate([p[1] for p in parts])
        rels[rid] = distinct(upload(first, second))  # the reference's BTreeSet: sorted unique

    def get(rid: int):
        return rels.get(rid) or empty()

    def derive(rule: DerivedRule):
        if isinstance(rule, ProjectRule):
            src = get(rule.source)
            return distinct(ex.Relation((src.columns[rule.columns[0]], src.columns[rule.columns[1]])))
        if isinstance(rule, JoinRule):
            left, right = get(rule.left), get(rule.right)
            if len(left) == 0 or len(right) == 0:
                return empty()
            index = ex.RangeIndex.build(right.columns[rule.join.right_key], stream)
            return distinct(ex.join_hip(left, right, index, rule.join, stream))
        pos, negative = get(rule.positive), get(rule.negative)
        keyed = distinct(ex.Relation((pos.columns[rule.key_columns[0]], pos.columns[rule.key_columns[1]])))
        kept = keyed if len(negative) == 0 or len(keyed) == 0 else ex.anti_join_hip(
            keyed, negative, SortedBinaryAntiJoin(), stream)
        return distinct(ex.Relation((kept.columns[rule.head_columns[0]], kept.columns[rule.head_columns[1]])))

    summaries = []
    store = None
    for step in plan.steps:
        if isinstance(step, DerivedStep):
            out = get(step.target)
            for rule in step.rules:
                out = union(out, derive(rule))
            rels[step.target] = out
            continue
        seeds = {int(r): get(int(r)) for r in step.plan.relations}
        for target, rule in step.seeds:
            seeds[target] = union(seeds[target], derive(rule))
        if store is None:
            store = RelationStore(stream)
        store.clear()
        members = {int(r) for r in step.plan.relations}
        for rid in {int(r.right_input) for r in step.plan.rules} - members:
            right = get(rid)
            store.insert_static(RelationId(rid), right.columns)
        for rid in members:
            s = seeds[rid]
            store.insert_recursive(RelationId(rid), s.columns, s.columns)
        summary = RecursiveExecutor.compile(step.plan, store).run(store, max_iterations)
        for rid in members:
            rels[rid] = ex.Relation(store.full_device(RelationId(rid)))
        summaries.append(SccRunSummary([(int(r), n) for r, n in summary.relation_rows], summary.iterations,
                                       summary.total_new_rows))
    if store is not None:
        store.close()
    return ProgramExecution(rels, summaries)


# ---------------------------------------------------------------- database

class ExecutionBackend(Enum):
    """database.rs:282-286 plus the device backend."""

    GENERAL_CPU = "general_cpu"
    BINARY_HYBRID = "binary_hybrid"
    BINARY_HIP = "binary_hip"


class DatabaseError(Exception):
    """database.rs:288-308."""


class ParseFailed(DatabaseError):
    def __init__(self, diagnostics):
        super().__init__("; ".join(d.message for d in diagnostics))
        self.diagnostics = diagnostics


class ResolveFailed(DatabaseError):
    def __init__(self, diagnostics):
        super().__init__("; ".join(d.message for d in diagnostics))
        self.diagnostics = diagnostics


class LoweringFailed(DatabaseError):
    def __init__(self, error: Exception):
        super().__init__(str(error))
        self.error = error


class NoProgram(DatabaseError):
    def __init__(self):
        super().__init__("no program loaded")


class NotRun(DatabaseError):
    def __init__(self):
        super().__init__("the program has not been run since it last changed")


class UnknownPredicate(DatabaseError):
    def __init__(self, name: str):
        super().__init__(f"unknown predicate {name}")
        self.predicate = name


class ArityMismatch(DatabaseError):
    def __init__(self, predicate: str, expected: int, actual: int):
        super().__init__(f"{predicate} expects {expected} values, got {actual}")
        self.predicate, self.expected, self.actual = predicate, expected, actual


class DelimitedError(DatabaseError):
    def __init__(self, line: int, column: int, message: str):
        super().__init__(f"line {line}, column {column}: {message}")
        self.line, self.column, self.message = line, column, message


@dataclass
class QueryResult:
    predicate: str
    rows: list[list]


@dataclass
class RunSummary:
    backend: ExecutionBackend
    sccs: list[SccRunSummary]


def parse_delimited(source: str, delimiter: str = ",", has_header: bool = False) -> list[list]:
    """parse_delimited_parallel (ingest.rs:20-120): quoted fields are
    strings, unquoted fields u32 when they parse, symbols otherwise."""
    if delimiter.isspace() and delimiter != "\t":
        raise DelimitedError(0, 0, "delimiter must be a visible ASCII byte or tab")
    out = []
    for lineno, record in enumerate(source.splitlines(), start=1):
        if has_header and lineno == 1:
            continue
        if not record.strip():
            continue
        fields, pos, n = [], 0, len(record)
        while True:
            column = pos + 1
            if pos < n and record[pos] == '"':
                pos += 1
                value, closed = [], False
                while pos < n:
                    if record[pos] == '"':
                        if pos + 1 < n and record[pos + 1] == '"':
                            value.append('"')
                            pos += 2
                        else:
                            pos += 1
                            closed = True
                            break
                    else:
                        value.append(record[pos])
                        pos += 1
                if not closed:
                    raise DelimitedError(lineno, column, "unterminated quoted field")
                fields.append(Text("".join(value)))
                if pos < n and record[pos] != delimiter:
                    raise DelimitedError(lineno, pos + 1, "expected delimiter after quoted field")
            else:
                start = pos
                while pos < n and record[pos] != delimiter:
                    pos += 1
                value = record[start:pos].strip()
                if not value:
                    raise DelimitedError(lineno, column, "empty fields are not values")
                number = int(value) if re.fullmatch(r"\+?[0-9]+", value) else None
                fields.append(number if number is not None and number <= U32_MAX else Symbol(value))
            if pos == n:
                break
            pos += 1
            if pos == n:
                raise DelimitedError(lineno, pos, "record ends with an empty field")
        out.append(fields)
    return out


class Database:
    """database.rs:19-290 on the device pipeline."""

    def __init__(self, stream=None, catalog: ProgramCatalog | None = None):
        self.catalog = catalog or ProgramCatalog()
        self.stream = stream
        self.program: ResolvedProgram | None = None
        self.plan: HipProgramPlan | None = None
        self.active: set[int] = set()
        self.inserted: dict[int, list[tuple[np.ndarray, np.ndarray]]] = {}
        self.execution: ProgramExecution | None = None

    def load_program(self, source: str) -> None:
        """database.rs:48-78."""
        parsed, diags = parse_program(source)
        if diags:
            raise ParseFailed(diags)
        resolved, diags = resolve_program(parsed, self.catalog)
        if diags:
            raise ResolveFailed(diags)
        try:
            plan = lower_binary_hip(resolved)
        except (LoweringError, StratificationError) as e:
            raise LoweringFailed(e) from e
        self.program, self.plan = resolved, plan
        self.active = ({r.head.predicate for r in resolved.rules} |
                       {l.atom.predicate for r in resolved.rules for l in r.body} | set(resolved.declarations))
        self.inserted = {}
        self.execution = None

    def _predicate(self, name: str) -> int:
        pid = self.catalog.predicates.id(name)
        if pid is None or pid not in self.active:
            raise UnknownPredicate(name)
        return pid

    def _insert_ids(self, pid: int, first: np.ndarray, second: np.ndarray) -> None:
        self.inserted.setdefault(pid, []).append((first, second))
        self.execution = None

    def insert(self, predicate: str, values) -> None:
        """database.rs:80-111: one fact, values interned (int = u32, str or
        Symbol = symbol, Text = string)."""
        pid = self._predicate(predicate)
        values = list(values)
        arity = self.catalog.predicates.get(pid).arity
        if len(values) != arity:
            raise ArityMismatch(predicate, arity, len(values))
        ids = [self.catalog.values.intern(v) for v in values]
        self._insert_ids(pid, np.array([ids[0]], np.uint32), np.array([ids[1]], np.uint32))

    def load_delimited(self, predicate: str, source: str, delimiter: str = ",", has_header: bool = False) -> int:
        """database.rs:156-168."""
        rows = parse_delimited(source, delimiter, has_header)
        for row in rows:
            self.insert(predicate, row)
        return len(rows)

    def load_binary_u32(self, predicate: str, path) -> int:
        """A whitespace-separated u32 file (storage/src/io.rs:72-141) parsed
        on the device, its values interned as u32 constants."""
        from . import ingest

        pid = self._predicate(predicate)
        arity = self.catalog.predicates.get(pid).arity
        if arity != 2:
            raise ArityMismatch(predicate, arity, 2)
        rel = ingest.load_binary_u32(path, self.stream)
        first, second = rel.to_host()
        a, b = self.catalog.values.intern_u32_rows(first, second)
        self._insert_ids(pid, a, b)
        return int(a.size)

    def run(self, max_iterations: int = 10_000) -> RunSummary:
        """database.rs:113-150 with the device backend."""
        if self.plan is None:
            raise NoProgram()
        extra = {pid: (np.concatenate([p[0] for p in parts]), np.concatenate([p[1] for p in parts]))
                 for pid, parts in self.inserted.items()}
        self.execution = execute_binary_hip(self.plan, self.stream, max_iterations, extra)
        return RunSummary(ExecutionBackend.BINARY_HIP, self.execution.summaries)

    def query(self, predicate: str) -> QueryResult:
        """database.rs:170-218: FULL in store order (sorted by interned ids),
        decoded."""
        pid = self._predicate(predicate)
        if self.execution is None:
            raise NotRun()
        rel = self.execution.relations.get(pid)
        if rel is None or len(rel) == 0:
            return QueryResult(predicate, [])
        first, second = rel.to_host()
        get = self.catalog.values.get
        return QueryResult(predicate, [[get(int(a)), get(int(b))] for a, b in zip(first, second)])

    def outputs(self) -> list[QueryResult]:
        if self.program is None:
            raise NoProgram()
        return [self.query(self.catalog.predicates.get(p).name) for p in self.program.outputs]

    def explain(self) -> str:
        if self.plan is None:
            raise NoProgram()
        return explain(self.plan, self.catalog)
