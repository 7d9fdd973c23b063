"""sparkalog_amd: MI355X-native binary recursive operator pipeline.

Modules:
  plans      relational plan vocabulary (BinaryEqualityJoin, ...)
  execution  operator API over device relations (join/distinct/anti-join/union/filter)
  recursion  device-resident RelationStore and RecursiveExecutor
  workload   seeded synthetic workloads
  build      in-tree build of the native libraries
"""
from .plans import (BinaryDistinct, BinaryEqualityJoin, JoinInput, JoinProjection, RecursiveRulePlan,
                    RecursiveSccPlan, RelationId, SortedBinaryAntiJoin, SortedBinaryUnion, U32Predicate,
                    path_join_plan, transitive_closure_scc)

__all__ = [
    "BinaryDistinct", "BinaryEqualityJoin", "JoinInput", "JoinProjection", "RecursiveRulePlan",
    "RecursiveSccPlan", "RelationId", "SortedBinaryAntiJoin", "SortedBinaryUnion", "U32Predicate",
    "path_join_plan", "transitive_closure_scc",
]
