"""Regenerate tests/golden/batch_expected.npz: per-instance closure facts of
the 4,096-instance batch (seed 0xC4), by brute-force BFS in the oracle.
Run from the repository root: python tests/golden/make_batch_fixture.py"""
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from sparkalog_amd import workload  # noqa: E402


def main():
    s, d, offs, counts = workload.batch_graphs()
    n = counts.size
    pairs = np.zeros(n, np.uint64)
    digest = np.zeros(n, np.uint64)
    max_dist = np.zeros(n, np.uint32)
    for i in range(n):
        r = oracle.tc_bfs(s[offs[i]:offs[i + 1]], d[offs[i]:offs[i + 1]], int(counts[i]), threads=8)
        pairs[i], digest[i], max_dist[i] = r["pairs"], r["digest"], r["max_dist"]
    out = ROOT / "tests" / "golden" / "batch_expected.npz"
    np.savez_compressed(out, node_counts=counts, edges=np.diff(offs).astype(np.uint64), pairs=pairs, digest=digest,
                        max_dist=max_dist, edge_digest=np.uint64(workload.pair_digest(s, d)))
    print("wrote", out, "total pairs", int(pairs.sum()), "max iterations", int(max_dist.max()))


if __name__ == "__main__":
    main()
