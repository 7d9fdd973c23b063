"""Regenerate tests/golden/workload_expected.json.

For each seeded workload this records the input digest and the closure facts
computed by brute force (per-source BFS in the oracle), and for the 1,000-node
DAG also by the general CPU path restatement. Run from the repository root:
    python tests/golden/make_fixtures.py
"""
import json
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from sparkalog_amd import workload  # noqa: E402


def closure_facts(src, dst, nodes):
    bfs = oracle.tc_bfs(src, dst, nodes, threads=8)
    return {"path_rows": int(bfs["pairs"]), "path_digest": int(bfs["digest"]), "max_dist": int(bfs["max_dist"])}


def main():
    out = {}
    # plain TC input: 1,000-node DAG with 4,000 distinct edges
    s, d = workload.random_dag()
    c1 = {"seed": workload.C1_SEED, "nodes": 1000, "edges": int(s.size),
          "edge_digest": int(workload.pair_digest(s, d))}
    c1.update(closure_facts(s, d, 1000))
    g = oracle.general_tc(s, d)
    assert g["path"][0].size == c1["path_rows"]
    assert int(workload.pair_digest(*g["path"])) == c1["path_digest"]
    assert g["iterations"] == c1["max_dist"]
    c1["nonlocal_rows"] = int(g["nonlocal"][0].size)
    c1["nonlocal_digest"] = int(workload.pair_digest(*g["nonlocal"]))
    c1["new_rows"] = int(g["new_rows"])
    out["random_dag_1000_4000"] = c1
    # single-instance TC: 2^20 nodes, Zipf(1.5) components, Poisson(2) out-degree
    for nodes in (1 << 14, 1 << 17, 1 << 20):
        s, d, comps = workload.component_graph(nodes=nodes)
        e = {"seed": workload.C2_SEED, "nodes": nodes, "components": comps, "edges": int(s.size),
             "edge_digest": int(workload.pair_digest(s, d))}
        e.update(closure_facts(s, d, nodes))
        out[f"component_graph_{nodes}"] = e
        print(nodes, e, flush=True)
    path = ROOT / "tests" / "golden" / "workload_expected.json"
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
