"""Relational plan vocabulary of the binary pipeline.

Mirrors crates/sparkalog-relational/src/lib.rs: the operator plans
(BinaryEqualityJoin :46-50, BinaryDistinct :54-56, SortedBinaryAntiJoin
:60-63, SortedBinaryUnion :67-70), the recursive plans (RecursiveRulePlan
:74-82, RecursiveSccPlan :86-89), U32Predicate (:144-164), and
transitive_closure_scc (crates/sparkalog-recursion/src/generic.rs:629-661).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from enum import IntEnum


class JoinInput(IntEnum):
    LEFT = 0
    RIGHT = 1


@dataclass(frozen=True)
class JoinProjection:
    input: JoinInput
    column: int


@dataclass(frozen=True)
class BinaryEqualityJoin:
    left_key: int
    right_key: int
    output: tuple[JoinProjection, JoinProjection]


@dataclass(frozen=True)
class BinaryDistinct:
    columns: tuple[int, int] = (0, 1)


@dataclass(frozen=True)
class SortedBinaryAntiJoin:
    left: tuple[int, int] = (0, 1)
    right: tuple[int, int] = (0, 1)


@dataclass(frozen=True)
class SortedBinaryUnion:
    left: tuple[int, int] = (0, 1)
    right: tuple[int, int] = (0, 1)


@dataclass(frozen=True)
class RelationId:
    value: int

    def __index__(self) -> int:
        return self.value


@dataclass(frozen=True)
class RecursiveRulePlan:
    target: RelationId
    delta_input: RelationId
    right_input: RelationId
    join: BinaryEqualityJoin
    distinct: BinaryDistinct = field(default_factory=BinaryDistinct)
    anti_join: SortedBinaryAntiJoin = field(default_factory=SortedBinaryAntiJoin)
    union: SortedBinaryUnion = field(default_factory=SortedBinaryUnion)


@dataclass(frozen=True)
class RecursiveSccPlan:
    relations: tuple[RelationId, ...]
    rules: tuple[RecursiveRulePlan, ...]


class U32Predicate:
    """A comparison over one u32 column; op codes as encode_predicate
    (crates/sparkalog-execution/src/lib.rs:1769-1778)."""

    EQ, NE, LT, LE, GT, GE = range(6)
    _NAMES = ("Eq", "Ne", "Lt", "Le", "Gt", "Ge")

    def __init__(self, op: int, value: int):
        if op not in range(6):
            raise ValueError(f"unknown predicate op {op}")
        self.op = op
        self.value = value & 0xFFFFFFFF

    @classmethod
    def eq(cls, v): return cls(cls.EQ, v)
    @classmethod
    def ne(cls, v): return cls(cls.NE, v)
    @classmethod
    def lt(cls, v): return cls(cls.LT, v)
    @classmethod
    def le(cls, v): return cls(cls.LE, v)
    @classmethod
    def gt(cls, v): return cls(cls.GT, v)
    @classmethod
    def ge(cls, v): return cls(cls.GE, v)

    def matches(self, value: int) -> bool:
        v, o = value & 0xFFFFFFFF, self.value
        return (v == o, v != o, v < o, v <= o, v > o, v >= o)[self.op]

    def __repr__(self) -> str:
        return f"U32Predicate.{self._NAMES[self.op]}({self.value})"


def path_join_plan() -> BinaryEqualityJoin:
    """path(x,z) :- path(x,y), edge(y,z): join path.1 = edge.0, output
    (path.0, edge.1)."""
    return BinaryEqualityJoin(
        left_key=1, right_key=0,
        output=(JoinProjection(JoinInput.LEFT, 0), JoinProjection(JoinInput.RIGHT, 1)))


def transitive_closure_scc(path: RelationId, edge: RelationId) -> RecursiveSccPlan:
    return RecursiveSccPlan(
        relations=(path,),
        rules=(RecursiveRulePlan(target=path, delta_input=path, right_input=edge,
                                 join=path_join_plan()),))
