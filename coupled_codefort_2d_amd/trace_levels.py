"""Per-level launch durations from a rocprofv3 kernel trace.

Groups the dispatches of the solver's kernels by (kernel, workgroups along x, grid y, workgroup
size) -- for the multigrid level kernels the grid identifies the level -- and prints launches,
average and total duration per group, largest total first.  Usage:

    python -m coupled_codefort_2d_amd.trace_levels TRACE.csv OUT.txt

TRACE.csv is the `run_kernel_trace.csv` of `rocprofv3 --kernel-trace -f csv`.
"""
from __future__ import annotations

import csv
import sys
from collections import defaultdict

PREFIXES = ("k_mg_", "k_mgs_", "k_vmg_", "k_matvec", "k_mdot", "k_maxpy", "k_pcd", "k_fold", "k_lin")


def level_stats(src):
    acc = defaultdict(lambda: [0, 0.0])
    for r in csv.DictReader(open(src)):
        name = r["Kernel_Name"].split("(")[0].replace("void ", "").replace("fugu::", "")
        if not any(p in name for p in PREFIXES):
            continue
        gx, gy, wx = int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]), int(r["Workgroup_Size_X"])
        key = (name, gx // max(wx, 1), gy, wx)
        acc[key][0] += 1
        acc[key][1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1000.0
    return sorted(acc.items(), key=lambda kv: -kv[1][1])


def main(src, out):
    with open(out, "w") as f:
        f.write("kernel  workgroups_x  grid_y  wg_size  launches  avg_us  total_ms\n")
        for (name, bx, gy, wx), (n, t) in level_stats(src):
            f.write(f"{name}  {bx}  {gy}  {wx}  {n}  {t / n:.2f}  {t / 1000:.2f}\n")


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
