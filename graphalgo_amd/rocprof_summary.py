"""Summaries of rocprofv3 CSV output (``rocprofv3 ... -f csv -d DIR -o NAME``).

    python -m graphalgo_amd.rocprof_summary stats DIR [--top N]
        per-kernel table from *_kernel_stats.csv (calls, total ms, average us, %)
    python -m graphalgo_amd.rocprof_summary pmc DIR [--match SUBSTR]
        per-kernel counter table from *_counter_collection.csv (dispatches,
        sum, average per dispatch; FETCH_SIZE / WRITE_SIZE are in KiB)
    python -m graphalgo_amd.rocprof_summary trace DIR --match SUBSTR
        per-dispatch duration distribution of the matching kernel from
        *_kernel_trace.csv

Prints markdown, so a GPU run can keep the few-KB summary and drop the raw
files.
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
import statistics
import sys
from collections import defaultdict


def _one(dirpath: str, suffix: str) -> str:
    hits = glob.glob(os.path.join(dirpath, "**", f"*{suffix}"), recursive=True)
    if not hits:
        sys.exit(f"no *{suffix} under {dirpath}")
    return hits[0]


def _short(name: str, width: int = 90) -> str:
    name = name.replace("\n", " ")
    return name if len(name) <= width else name[: width - 3] + "..."


def stats(dirpath: str, top: int) -> None:
    rows = list(csv.DictReader(open(_one(dirpath, "_kernel_stats.csv"))))
    print("| kernel | calls | total ms | average us | % |")
    print("|---|---:|---:|---:|---:|")
    for r in rows[:top]:
        print(f"| `{_short(r['Name'])}` | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.2f} | "
              f"{float(r['AverageNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")


def pmc(dirpath: str, match: str | None) -> None:
    agg: dict = defaultdict(lambda: defaultdict(lambda: [0, 0.0]))
    for r in csv.DictReader(open(_one(dirpath, "_counter_collection.csv"))):
        if match and match not in r["Kernel_Name"]:
            continue
        a = agg[r["Kernel_Name"]][r["Counter_Name"]]
        a[0] += 1
        a[1] += float(r["Counter_Value"])
    print("| kernel | counter | dispatches | sum | average per dispatch |")
    print("|---|---|---:|---:|---:|")
    order = sorted(agg.items(), key=lambda kv: -max(v[1] for v in kv[1].values()))
    for name, counters in order[:12]:
        for cname, (n, total) in sorted(counters.items()):
            print(f"| `{_short(name, 70)}` | {cname} | {n} | {total:.1f} | {total / max(n, 1):.2f} |")


def trace(dirpath: str, match: str) -> None:
    d = []
    for r in csv.DictReader(open(_one(dirpath, "_kernel_trace.csv"))):
        if match in r["Kernel_Name"]:
            d.append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    if not d:
        sys.exit("no matching dispatches")
    d.sort()
    q = lambda p: d[min(len(d) - 1, int(p * len(d)))]  # noqa: E731
    print(f"dispatches {len(d)}, total {sum(d) / 1e3:.2f} ms, mean {statistics.mean(d):.2f} us, "
          f"p10 {q(0.1):.2f} p50 {q(0.5):.2f} p90 {q(0.9):.2f} max {d[-1]:.2f} us")


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["stats", "pmc", "trace"])
    ap.add_argument("dir")
    ap.add_argument("--top", type=int, default=15)
    ap.add_argument("--match", default=None)
    a = ap.parse_args()
    if a.mode == "stats":
        stats(a.dir, a.top)
    elif a.mode == "pmc":
        pmc(a.dir, a.match)
    else:
        trace(a.dir, a.match or "k_mxm_items")


if __name__ == "__main__":
    main()
