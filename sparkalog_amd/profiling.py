"""Summaries of rocprofv3 output for the kernels of the fixpoint.

Two summaries, both plain text so they can sit beside the code:

* ``stats``: the ``--kernel-trace --stats`` table (``*_kernel_stats.csv``)
  as markdown, one row per kernel with calls, total and average duration;
* ``traffic``: HBM bytes per dispatch from two ``--pmc`` runs (FETCH_SIZE,
  then WRITE_SIZE; both in KiB per dispatch), each scaled by what the same
  counter reads for a plain device copy of known size, written as JSON
  keyed by kernel name (the form ``bench.py`` reads for its roofline).

    python -m sparkalog_amd.profiling stats STATS_CSV OUT_MD --title T --command C
    python -m sparkalog_amd.profiling traffic --fetch CSV --write CSV \\
        --copy-fetch CSV --copy-write CSV --copy-bytes N OUT_JSON --command C
"""
from __future__ import annotations

import argparse
import csv
import json
from collections import defaultdict
from pathlib import Path


def short_name(name: str) -> str:
    """Kernel name without return type, namespaces, template arguments and
    parameter list: "void ns::(anonymous namespace)::k<int>(int*)" -> "k"."""
    name = name.strip().replace("(anonymous namespace)::", "")
    if name.endswith(")"):  # drop the trailing parameter list
        depth = 0
        for i in range(len(name) - 1, -1, -1):
            depth += {")": 1, "(": -1}.get(name[i], 0)
            if depth == 0:
                name = name[:i]
                break
    out, depth = [], 0  # drop template arguments
    for ch in name:
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        elif depth == 0:
            out.append(ch)
    base = "".join(out).split(" ")[-1]
    return base.rsplit("::", 1)[-1]


def templated_name(name: str) -> str:
    """short_name with the template arguments kept ("k<unsigned int, true>"),
    to tell instantiations of one kernel apart."""
    base = short_name(name)
    name = name.strip().replace("(anonymous namespace)::", "")
    at = name.find(base + "<")
    if at < 0:
        return base
    depth = 0
    for i in range(at + len(base), len(name)):
        depth += {"<": 1, ">": -1}.get(name[i], 0)
        if depth == 0:
            return name[at:i + 1]
    return base


def stats_markdown(stats_csv: Path, title: str, command: str, top: int = 40) -> str:
    rows = list(csv.DictReader(open(stats_csv)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    seen = defaultdict(int)
    for r in rows:
        seen[short_name(r["Name"])] += 1
    # instantiations of one template are listed with their arguments
    label = {r["Name"]: templated_name(r["Name"]) if seen[short_name(r["Name"])] > 1 else short_name(r["Name"])
             for r in rows}
    out = [f"# {title}", "", "Command (run from the repository root):", "", "```", command, "```", "",
           f"Total kernel time: {total / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} dispatches", "",
           "| Kernel | Calls | Total ms | Avg us | Min us | Max us | % |", "|---|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"]))[:top]:
        out.append(f"| {label[r['Name']]} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} | "
                   f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} | "
                   f"{float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out) + "\n"


def counter_per_kernel(counter_csv: Path, counter: str) -> dict[str, tuple[int, float]]:
    """kernel -> (dispatches, summed counter value) for one counter."""
    per_dispatch: dict[tuple[str, str], float] = defaultdict(float)
    for r in csv.DictReader(open(counter_csv)):
        if r["Counter_Name"] != counter:
            continue
        # one row per dispatch and counter (values of several instances are
        # already summed by the profiler for these derived counters)
        per_dispatch[(short_name(r["Kernel_Name"]), r["Dispatch_Id"])] += float(r["Counter_Value"])
    out: dict[str, list] = defaultdict(lambda: [0, 0.0])
    for (kernel, _), v in per_dispatch.items():
        out[kernel][0] += 1
        out[kernel][1] += v
    return {k: (int(v[0]), float(v[1])) for k, v in out.items()}


COPY_KERNEL = "copy_u64_kernel"


def copy_scale(counter_csv: Path, counter: str, copy_bytes: int) -> tuple[float, str]:
    """Bytes per counter unit: the copy's known bytes (read once, written
    once) over the counter's mean reading for the copy kernel."""
    n, total = counter_per_kernel(counter_csv, counter)[COPY_KERNEL]
    return copy_bytes / (total / n), COPY_KERNEL


def traffic_json(fetch_csv: Path, write_csv: Path, copy_fetch: Path, copy_write: Path, copy_bytes: int,
                 command: str) -> dict:
    fs, fk = copy_scale(copy_fetch, "FETCH_SIZE", copy_bytes)
    ws, wk = copy_scale(copy_write, "WRITE_SIZE", copy_bytes)
    fetch = counter_per_kernel(fetch_csv, "FETCH_SIZE")
    write = counter_per_kernel(write_csv, "WRITE_SIZE")
    per_launch, detail = {}, {}
    for kernel in sorted(set(fetch) | set(write)):
        fn, fv = fetch.get(kernel, (0, 0.0))
        wn, wv = write.get(kernel, (0, 0.0))
        fb = fv / fn * fs if fn else 0.0
        wb = wv / wn * ws if wn else 0.0
        per_launch[kernel] = round(fb + wb, 1)
        detail[kernel] = {"dispatches": max(fn, wn), "fetch_bytes_per_dispatch": round(fb, 1),
                          "write_bytes_per_dispatch": round(wb, 1)}
    return {"command": command, "copy_bytes": copy_bytes,
            "fetch_scale_bytes_per_unit": fs, "write_scale_bytes_per_unit": ws,
            "copy_kernels": {"fetch": fk, "write": wk},
            "traffic_bytes_per_launch": per_launch, "detail": detail}


def main(argv=None) -> None:
    p = argparse.ArgumentParser(prog="python -m sparkalog_amd.profiling")
    sub = p.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("stats_csv", type=Path)
    s.add_argument("out", type=Path)
    s.add_argument("--title", required=True)
    s.add_argument("--command", required=True)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", type=Path, required=True)
    t.add_argument("--write", type=Path, required=True)
    t.add_argument("--copy-fetch", type=Path, required=True)
    t.add_argument("--copy-write", type=Path, required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("out", type=Path)
    t.add_argument("--command", required=True)
    a = p.parse_args(argv)
    if a.what == "stats":
        a.out.write_text(stats_markdown(a.stats_csv, a.title, a.command))
    else:
        data = traffic_json(a.fetch, a.write, a.copy_fetch, a.copy_write, a.copy_bytes, a.command)
        a.out.write_text(json.dumps(data, indent=1) + "\n")


if __name__ == "__main__":
    main()
