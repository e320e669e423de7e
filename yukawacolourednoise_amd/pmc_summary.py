"""Summarise a rocprofv3 ``--pmc`` counter-collection CSV for chosen kernels.

    python3 -m yukawacolourednoise_amd.pmc_summary <counter_collection.csv> <name-part> [...]

For every kernel whose name contains one of the given parts, prints the number of
dispatches and, per counter, the median value over the dispatches that did work (those
whose SQ_WAVES, when collected, is at least half the largest seen for that kernel; the CG
solvers launch some chunks whose kernels exit at once for converged systems).  Values are
per dispatch, summed over the GPU's instances as rocprofv3 reports them."""
from __future__ import annotations

import csv
import statistics
import sys
from collections import defaultdict


def summarise(path: str, parts: list[str]) -> dict:
    # (kernel, dispatch) -> {counter: value}
    disp: dict = defaultdict(dict)
    for row in csv.DictReader(open(path, newline="")):
        name = row.get("Kernel_Name", "")
        if not any(p in name for p in parts):
            continue
        key = (name, row.get("Dispatch_Id") or row.get("Correlation_Id"))
        disp[key][row["Counter_Name"]] = disp[key].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
    per_kernel: dict = defaultdict(list)
    for (name, _), vals in disp.items():
        per_kernel[name].append(vals)
    out = {}
    for name, rows in per_kernel.items():
        if any("SQ_WAVES" in r for r in rows):
            top = max(r.get("SQ_WAVES", 0.0) for r in rows)
            rows = [r for r in rows if r.get("SQ_WAVES", 0.0) >= 0.5 * top]
        counters = sorted({c for r in rows for c in r})
        out[name] = {"dispatches": len(rows),
                     "median": {c: statistics.median(r[c] for r in rows if c in r) for c in counters}}
    return out


def main(argv: list[str]) -> int:
    if len(argv) < 3:
        print(__doc__)
        return 2
    for name, s in summarise(argv[1], argv[2:]).items():
        print(f"{name}\n  working dispatches: {s['dispatches']}")
        for c, v in s["median"].items():
            print(f"  {c:28s} {v:16.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
