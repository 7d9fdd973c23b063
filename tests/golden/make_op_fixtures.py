"""Regenerate tests/golden/op_expected.json: the operator microbench cases at
2^26 and 2^28 rows (sparkalog_amd.opbench.host_inputs, keys uniform and
Zipf(1.1)), each run through the oracle (the CPU restatement of the
reference's operators) on this host's cores.

For every case it records the input's digest (pinning the generator), the
output row count, the order-independent digest of the output pairs and, for
the join (whose output is not sorted), the order-dependent digest of
tests/helpers.py. Run from the repository root:
    python tests/golden/make_op_fixtures.py [log2_rows ...]
"""
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from helpers import operator_input_digest, oracle_operator, ordered_digest, strictly_sorted  # noqa: E402
from sparkalog_amd import opbench, workload  # noqa: E402

PATH = ROOT / "tests" / "golden" / "op_expected.json"
MODES = {"uniform": workload.KEY_UNIFORM, "zipf1.1": workload.KEY_ZIPF}


def main():
    sizes = [int(x) for x in sys.argv[1:]] or [26, 28]
    out = json.loads(PATH.read_text()) if PATH.exists() else {}
    threads = os.cpu_count() or 1
    for k in sizes:
        for name in opbench.OPERATORS:
            for mode_name, mode in MODES.items():
                t0 = time.time()
                host = opbench.host_inputs(name, 1 << k, mode)
                res = oracle_operator(name, host, threads)
                e = {"rows": 1 << k, "input_digest": operator_input_digest(name, host), "out_rows": int(res[0].size),
                     "digest": int(workload.pair_digest(*res))}
                if name == "join":
                    e["ordered_digest"] = ordered_digest(*res)
                else:
                    assert strictly_sorted(*res)
                out[f"{name}_2p{k}_{mode_name}"] = e
                print(name, k, mode_name, e, f"{time.time() - t0:.1f} s", flush=True)
                PATH.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
                del host, res


if __name__ == "__main__":
    main()
