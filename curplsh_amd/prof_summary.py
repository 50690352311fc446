"""Summarise a rocprofv3 kernel trace (``--kernel-trace --output-format csv``)
per (kernel, grid size): dispatch count and average / min / max duration.

    python -m curplsh_amd.prof_summary <dir or *_kernel_trace.csv> [--match SUBSTR] [--min-ms MS] [--last N]

A filter launch over the whole database has its own grid size, so its row is
the per-launch figure to compare with bench.py's ``roofline.ms_per_launch``
(HIP events around the same launch).
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os


def load(path: str):
    if os.path.isdir(path):
        files = glob.glob(os.path.join(path, "**", "*kernel_trace.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no *kernel_trace.csv under {path}")
        path = files[0]
    with open(path) as fh:
        return list(csv.DictReader(fh))


def summarize(rows, match: str = "", min_ms: float = 0.0, last: int = 0):
    groups = collections.defaultdict(list)
    rows = sorted(rows, key=lambda r: int(r["Start_Timestamp"]))
    for r in rows:
        name = r["Kernel_Name"]
        if match and match not in name:
            continue
        dur = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6
        if dur < min_ms:
            continue
        grid = r.get("Grid_Size") or r.get("Grid_Size_X") or "?"
        groups[(name, grid)].append(dur)
    out = []
    for (name, grid), d in groups.items():
        if last > 0:
            d = d[-last:]
        out.append((name, grid, len(d), sum(d) / len(d), min(d), max(d)))
    out.sort(key=lambda t: -t[2] * t[3])
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("trace")
    ap.add_argument("--match", default="")
    ap.add_argument("--min-ms", type=float, default=0.0)
    ap.add_argument("--top", type=int, default=20)
    ap.add_argument("--last", type=int, default=0,
                    help="only the last N dispatches of each (kernel, grid): a benchmark's timed steps")
    a = ap.parse_args()
    rows = summarize(load(a.trace), a.match, a.min_ms, a.last)
    print("| kernel | grid (threads) | dispatches | avg ms | min ms | max ms |")
    print("|---|---|---|---|---|---|")
    for name, grid, n, avg, lo, hi in rows[: a.top]:
        short = name.replace("void ", "").replace("rplsh::(anonymous namespace)::", "")
        short = short.replace("(rplsh::Elem)", "E").replace("(rplsh::FilterArgs)", "")
        short = short.replace("(rplsh::SelectArgs)", "").replace("(rplsh::RerankArgs)", "")[:100]
        print(f"| `{short}` | {grid} | {n} | {avg:.3f} | {lo:.3f} | {hi:.3f} |")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
