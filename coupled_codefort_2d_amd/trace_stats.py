"""Kernel statistics of a rocprofv3 run as a markdown table (top N kernels by total time).

Reads the `run_kernel_stats.csv` of `rocprofv3 --kernel-trace --stats -f csv`, merges the rows of
one kernel (template arguments kept, call arguments dropped) and prints calls, total and average
duration and share.  Usage:

    python -m coupled_codefort_2d_amd.trace_stats STATS.csv TITLE COMMAND [N]
"""
from __future__ import annotations

import csv
import re
import sys


def clean(name: str) -> str:
    name = re.sub(r"^void ", "", name)
    name = name.replace("fugu::", "").replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*\)$", "", name).strip()


def main(argv):
    path, title, cmd = argv[1], argv[2], argv[3]
    n = int(argv[4]) if len(argv) > 4 else 25
    agg = {}
    for r in csv.DictReader(open(path)):
        a = agg.setdefault(clean(r["Name"]), [0, 0.0])
        a[0] += int(r["Calls"])
        a[1] += float(r["TotalDurationNs"])
    tot = sum(v[1] for v in agg.values())
    calls = sum(v[0] for v in agg.values())
    print(f"# {title}\n\nCommand (from /tmp with TMPDIR=/tmp; runs committed files only):\n\n```\n{cmd}\n```\n")
    print(f"Total kernel time: {tot / 1e6:.1f} ms over {calls} launches ({len(agg)} kernels). Durations are "
          f"under the profiler. Top {n} by total time:\n")
    print("| kernel | calls | total ms | avg us | % |\n|---|---|---|---|---|")
    for k, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1])[:n]:
        print(f"| `{k}` | {c} | {t / 1e6:.2f} | {t / c / 1e3:.2f} | {100 * t / tot:.2f} |")


if __name__ == "__main__":
    main(sys.argv)
