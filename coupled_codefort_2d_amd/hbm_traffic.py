"""HBM traffic per launch from rocprofv3 counter runs.

Pairs the per-dispatch FETCH_SIZE and WRITE_SIZE values (KiB, one `--pmc` run each, CSV
output) of one kernel with the library's launch log (FUGU_LAUNCH_LOG=<path>: one line
"name algorithmic_bytes timed" per launch of a timed kernel, in launch order) and with a plain
copy kernel of known size from the same runs (bench.py --copy-probe-mib), whose readings give
the counter scale factors:

    fetch_factor = copy bytes read    / (FETCH_SIZE of the copy * 1024)
    write_factor = copy bytes written / (WRITE_SIZE of the copy * 1024)

KERNEL is the prefix of the device kernel name as rocprofv3 prints it after the namespace
(e.g. "k_mg_pre<double," or "k_maxpy_dot<double,"), NAME the library's timing name of the same
launches (e.g. "mg_pre_f64").  Prints the measured bytes per timed launch, their algorithmic
bytes and the ratio (traffic / algorithmic).  Usage:

    python -m coupled_codefort_2d_amd.hbm_traffic KERNEL NAME COPY_MIB FETCH.csv WRITE.csv LAUNCH.log
"""
from __future__ import annotations

import csv
import sys


def _dispatches(path, kernel):
    """Counter values (KiB) of the kernels named `kernel`... in dispatch order, and those of the
    copy probe."""
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Dispatch_Id"]))
    mine, copy = [], []
    for r in rows:
        name = r["Kernel_Name"].split("(")[0].split("::")[-1]
        if name.startswith(kernel):
            mine.append(float(r["Counter_Value"]))
        elif "MulFunctor" in r["Kernel_Name"]:
            copy.append(float(r["Counter_Value"]))
    return mine, copy


def traffic(kernel, name, copy_mib, fetch_csv, write_csv, launch_log):
    fetch, fcopy = _dispatches(fetch_csv, kernel)
    write, wcopy = _dispatches(write_csv, kernel)
    log = [ln.split() for ln in open(launch_log) if ln.split() and ln.split()[0] == name]
    if not (len(fetch) == len(write) == len(log)):
        raise ValueError(f"dispatch counts differ: fetch {len(fetch)} write {len(write)} log {len(log)}")
    copy_bytes = copy_mib * (1 << 20)
    ff = copy_bytes / (max(fcopy) * 1024.0)
    wf = copy_bytes / (max(wcopy) * 1024.0)
    meas = alg = 0.0
    n = 0
    for f, w, (_, ab, timed) in zip(fetch, write, log):
        if timed == "1":
            meas += f * 1024.0 * ff + w * 1024.0 * wf
            alg += float(ab)
            n += 1
    return {"kernel": kernel, "name": name, "launches": n, "fetch_factor": ff, "write_factor": wf,
            "traffic_per_launch": meas / n, "algorithmic_per_launch": alg / n, "ratio": meas / alg}


if __name__ == "__main__":
    k, nm, mib, fcsv, wcsv, lg = sys.argv[1:7]
    r = traffic(k, nm, int(mib), fcsv, wcsv, lg)
    for key, val in r.items():
        print(f"{key}: {val:.6g}" if isinstance(val, float) else f"{key}: {val}")
