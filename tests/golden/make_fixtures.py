"""Regenerate tests/golden/workload_expected.json.

For each seeded workload this records the input digest and the closure facts
computed by brute force (per-source BFS in the oracle), and for the 1,000-node
DAG also by the general CPU path restatement. Run from the repository root:
    python tests/golden/make_fixtures.py
(`python tests/golden/make_fixtures.py wide` refreshes only the 2^17-node
graph relabelled over the whole u32 range.)
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


def wide_facts(nodes):
    """The component graph with node ids relabelled over the whole u32 range
    (workload.wide_labels): the closure has the same rows and longest
    shortest path; its digest is taken over the relabelled pairs."""
    s, d, comps = workload.component_graph(nodes=nodes)
    labels = workload.wide_labels(nodes)
    ws, wd = workload.relabel(s, d, labels)
    bfs = oracle.tc_bfs(s, d, nodes, threads=8, closure=True)
    a, b = bfs["closure"]
    digest = 0
    step = 1 << 24
    for i in range(0, a.size, step):
        digest = (digest + int(workload.pair_digest(labels[a[i:i + step]], labels[b[i:i + step]]))) % (1 << 64)
    return {"seed": workload.C2_SEED, "nodes": nodes, "components": comps, "edges": int(ws.size),
            "labels": "wide_labels(nodes, 0xF011)", "edge_digest": int(workload.pair_digest(ws, wd)),
            "path_rows": int(bfs["pairs"]), "path_digest": digest, "max_dist": int(bfs["max_dist"])}


def main():
    path = ROOT / "tests" / "golden" / "workload_expected.json"
    if sys.argv[1:] == ["wide"]:
        out = json.loads(path.read_text())
        out["component_graph_131072_wide"] = wide_facts(1 << 17)
        print(out["component_graph_131072_wide"], flush=True)
        path.write_text(json.dumps(out, indent=1) + "\n")
        return
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
    out["component_graph_131072_wide"] = wide_facts(1 << 17)
    path.write_text(json.dumps(out, indent=1) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
