"""Stratified binary programs on the device.

The reference lowers a Datalog program to the binary pipeline only when
every derived predicate is recursive (`lower_binary`,
crates/sparkalog-datalog/src/subset.rs:39-97, rejection at :56-63); a program
with a non-recursive rule such as

    path(x, y) :- edge(x, y).
    path(x, z) :- path(x, y), edge(y, z).
    nonlocal(x, z) :- path(x, z), !edge(x, z).

falls back to the general CPU evaluator (crates/sparkalog-datalog/src/
database.rs:114-121), whose negation reads FULL of an earlier stratum
(general.rs:522-536). Here the same program runs in strata on one device
store: each recursive SCC with the semi-naive executor (as `execute_binary`,
subset.rs:369-414, does per SCC), then each negation rule as one sorted
anti-join of two stored relations, written as a static relation for the
strata after it.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Union

from .plans import RecursiveSccPlan, RelationId
from .recursion import RecursiveExecutor, RelationStore, SccSummary


@dataclass(frozen=True)
class NegationRulePlan:
    """target(x, y) :- positive(x, y), !negative(x, y)."""

    target: RelationId
    positive: RelationId
    negative: RelationId


Stratum = Union[RecursiveSccPlan, NegationRulePlan]


@dataclass(frozen=True)
class BinaryProgramPlan:
    """Strata in evaluation order (a stratum reads only relations stored
    before it)."""

    strata: tuple[Stratum, ...]


@dataclass
class ProgramSummary:
    scc: list[SccSummary]
    derived_rows: dict[int, int]


def run_program(plan: BinaryProgramPlan, store: RelationStore, max_iterations: int = 10_000) -> ProgramSummary:
    """Evaluate every stratum in order on `store`; the relations a stratum
    reads must be stored already (inserted, or derived by an earlier stratum)."""
    sccs: list[SccSummary] = []
    derived: dict[int, int] = {}
    for stratum in plan.strata:
        if isinstance(stratum, RecursiveSccPlan):
            summary = RecursiveExecutor.compile(stratum, store).run(store, max_iterations)
            sccs.append(summary)
            for rid, rows in summary.relation_rows:
                derived[int(rid)] = rows
        elif isinstance(stratum, NegationRulePlan):
            derived[int(stratum.target)] = store.difference(stratum.target, stratum.positive, stratum.negative)
        else:
            raise TypeError(f"unknown stratum {stratum!r}")
    return ProgramSummary(sccs, derived)


def transitive_closure_nonlocal(path: RelationId, edge: RelationId, nonlocal_: RelationId) -> BinaryProgramPlan:
    """path = TC(edge); nonlocal = path \\ edge."""
    from .plans import transitive_closure_scc

    return BinaryProgramPlan((transitive_closure_scc(path, edge), NegationRulePlan(nonlocal_, path, edge)))
