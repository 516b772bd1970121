"""Short text summaries of rocprofv3 CSV output, for the committed profile summaries.

    python -m gridcl_amd.profsum stats  <prefix>_kernel_stats.csv [--top N]
    python -m gridcl_amd.profsum pmc    <prefix>_counter_collection.csv [--match SUBSTRING] [--runs]
    python -m gridcl_amd.profsum trace  <prefix>_kernel_trace.csv [--match SUBSTRING] [--last N]
                                        [--runs]
    python -m gridcl_amd.profsum digest <kernel source> [...]

``stats`` prints calls and average/min/max duration per kernel; ``pmc`` prints, per kernel,
counter and grid size, the number of dispatches and the mean, min and max counter value per
dispatch (FETCH_SIZE and WRITE_SIZE are in KiB); ``trace`` prints, per kernel and grid size, the
dispatch count and the average, min and max duration, optionally of the last N dispatches only
(the timed steps of a run); with ``--runs`` both split each kernel and grid size into runs of
back-to-back dispatches (a dispatch of any other kernel or grid ends a run), numbered from 0 in
dispatch order, so that legs of one program that launch the same kernel at the same grid (one
point set after another) are told apart; ``digest`` prints the content digest of kernel sources and the local
headers they include, which each profile summary records so that a reader (and the benchmark,
for the traffic figure) can tell whether the profiled kernel is still the one in the tree.
"""
from __future__ import annotations

import argparse
import csv
import hashlib
import os
import re
import sys
from collections import defaultdict

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")


def source_digest(*names: str) -> str:
    """The first 16 hex digits of SHA-256 over the named csrc files and the local headers they
    include, transitively, in sorted order (name and content)."""
    seen, todo = set(), list(names)
    while todo:
        n = todo.pop()
        if n in seen:
            continue
        seen.add(n)
        with open(os.path.join(CSRC, n), encoding="utf-8") as f:
            todo += re.findall(r'^#include "([^"]+)"', f.read(), flags=re.M)
    h = hashlib.sha256()
    for n in sorted(seen):
        h.update(n.encode() + b"\0")
        with open(os.path.join(CSRC, n), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def short_name(name: str, width: int = 96) -> str:
    name = name.replace("tg::(anonymous namespace)::", "").replace("tg::", "")
    name = re.sub(r"\(.*\)$", "", name)  # drop the parameter list
    return name if len(name) <= width else name[: width - 3] + "..."


def stats(path: str, top: int) -> None:
    rows = list(csv.DictReader(open(path)))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print(f"{'kernel':<72} {'calls':>6} {'avg us':>9} {'min us':>9} {'max us':>9} {'%':>6}")
    for r in rows[:top]:
        print(f"{short_name(r['Name'], 72):<72} {int(r['Calls']):>6} "
              f"{float(r['AverageNs']) / 1e3:>9.1f} {float(r['MinNs']) / 1e3:>9.1f} "
              f"{float(r['MaxNs']) / 1e3:>9.1f} {float(r['Percentage']):>6.2f}")


def run_numbers(rows, key) -> dict:
    """Dispatch id -> the number of its run among the runs of its key (kernel and grid): the
    rows in dispatch order, a run ending where a dispatch of another key comes between."""
    order = sorted({(int(r["Dispatch_Id"]), key(r)) for r in rows})
    runs, out, prev = defaultdict(int), {}, None
    for did, k in order:
        if prev is not None and k != prev and k in runs:
            runs[k] += 1
        runs.setdefault(k, 0)
        out[did] = runs[k]
        prev = k
    return out


def pmc(path: str, match: str | None, drop: int = 0, by_run: bool = False) -> None:
    rows = list(csv.DictReader(open(path)))
    run = run_numbers(rows, lambda r: (r["Kernel_Name"], r["Grid_Size"])) if by_run else {}
    acc = defaultdict(list)  # (kernel, counter, grid size in work-items) -> values
    for r in rows:
        if match and match not in r["Kernel_Name"]:
            continue
        name = short_name(r["Kernel_Name"], 72 - 6 * by_run)
        if by_run:
            name = f"{name} run {run[int(r['Dispatch_Id'])]}"
        key = (name, r["Counter_Name"], r["Grid_Size"])
        acc[key].append(float(r["Counter_Value"]))
    print(f"{'kernel':<72} {'counter':<11} {'n':>5} {'grid':>9} {'mean':>14} {'min':>14} "
          f"{'max':>14}")
    for (name, counter, grid), v in sorted(acc.items(), key=lambda kv: -max(kv[1])):
        if drop:
            v = v[:-drop] or v
        print(f"{name:<72} {counter:<11} {len(v):>5} {grid:>9} "
              f"{sum(v) / len(v):>14.1f} {min(v):>14.1f} {max(v):>14.1f}")


def trace(path: str, match: str | None, last: int, drop: int = 0, by_run: bool = False) -> None:
    rows = list(csv.DictReader(open(path)))
    grid_of = lambda r: int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
    run = run_numbers(rows, lambda r: (r["Kernel_Name"], grid_of(r))) if by_run else {}
    rows.sort(key=lambda r: int(r["Dispatch_Id"]))
    acc = defaultdict(list)  # (kernel, grid size in work-items) -> durations in ns
    for r in rows:
        if match and match not in r["Kernel_Name"]:
            continue
        name = short_name(r["Kernel_Name"], 72 - 6 * by_run)
        if by_run:
            name = f"{name} run {run[int(r['Dispatch_Id'])]}"
        acc[(name, grid_of(r))].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    print(f"{'kernel':<72} {'grid':>10} {'n':>6} {'avg us':>9} {'min us':>9} {'max us':>9}")
    for (name, grid), v in sorted(acc.items(), key=lambda kv: -sum(kv[1])):
        if drop:
            v = v[:-drop] or v
        if last:
            v = v[-last:]
        print(f"{name:<72} {grid:>10} {len(v):>6} {sum(v) / len(v) / 1e3:>9.1f} "
              f"{min(v) / 1e3:>9.1f} {max(v) / 1e3:>9.1f}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=12)
    b = sub.add_parser("pmc")
    b.add_argument("csv")
    b.add_argument("--match")
    b.add_argument("--drop", type=int, default=0,
                   help="leave out the last DROP dispatches of each (kernel, counter, grid)")
    b.add_argument("--runs", action="store_true",
                   help="split each (kernel, grid) into runs of back-to-back dispatches")
    c = sub.add_parser("trace")
    c.add_argument("csv")
    c.add_argument("--match")
    c.add_argument("--last", type=int, default=0)
    c.add_argument("--drop", type=int, default=0,
                   help="leave out the last DROP dispatches of each (kernel, grid) first")
    c.add_argument("--runs", action="store_true",
                   help="split each (kernel, grid) into runs of back-to-back dispatches")
    d = sub.add_parser("digest")
    d.add_argument("sources", nargs="+")
    args = ap.parse_args(argv)
    if args.cmd == "digest":
        print(source_digest(*args.sources))
    elif args.cmd == "stats":
        stats(args.csv, args.top)
    elif args.cmd == "pmc":
        pmc(args.csv, args.match, args.drop, args.runs)
    else:
        trace(args.csv, args.match, args.last, args.drop, args.runs)
    return 0


if __name__ == "__main__":
    sys.exit(main())
