"""Summarise rocprofv3 output into per-kernel tables.

    python -m nrpy_amd.profsummary <dir>/run_results.db      (kernel trace)
    python -m nrpy_amd.profsummary <dir>/run_counter_collection.csv   (--pmc, csv)

where <dir> is the rocprofv3 output directory (``-d``) and ``run`` its output name (``-o``).
"""

from __future__ import annotations

import sqlite3
import sys


def kernel_stats(db_path: str):
    db = sqlite3.connect(db_path)
    q = """select s.kernel_name, count(*), avg(d.end - d.start), min(d.end - d.start), max(d.end - d.start),
                  sum(d.end - d.start), s.arch_vgpr_count, s.accum_vgpr_count, s.private_segment_size
           from rocpd_kernel_dispatch d join rocpd_info_kernel_symbol s on d.kernel_id = s.id
           group by s.kernel_name order by 6 desc"""
    return list(db.execute(q))


def main() -> None:
    if sys.argv[1].endswith(".csv"):
        print_counters(sys.argv[1])
        return
    rows = kernel_stats(sys.argv[1])
    total = sum(r[5] for r in rows)
    print("| kernel | calls | avg us | min us | max us | total ms | % | arch VGPR | AGPR | scratch B/lane |")
    print("|---|---|---|---|---|---|---|---|---|---|")
    for name, n, avg, mn, mx, tot, v, a, scr in rows:
        short = name.replace(".kd", "")
        if len(short) > 60:
            short = short[:57] + "..."
        print(f"| {short} | {n} | {avg/1e3:.1f} | {mn/1e3:.1f} | {mx/1e3:.1f} | {tot/1e6:.3f} | "
              f"{100*tot/total:.1f} | {v} | {a} | {scr} |")


def counter_stats(csv_path: str):
    """Per kernel and counter: dispatches and mean value per dispatch, from a
    rocprofv3 ``--pmc ... -f csv`` counter_collection file."""
    import csv
    from collections import defaultdict

    acc = defaultdict(lambda: [0, 0.0])
    with open(csv_path, newline="") as f:
        for row in csv.DictReader(f):
            key = (row["Kernel_Name"], row["Counter_Name"])
            acc[key][0] += 1
            acc[key][1] += float(row["Counter_Value"])
    return sorted((k[0], k[1], n, s / n) for k, (n, s) in acc.items())


def print_counters(csv_path: str) -> None:
    print("| kernel | counter | dispatches | mean per dispatch |")
    print("|---|---|---|---|")
    for name, ctr, n, mean in counter_stats(csv_path):
        short = name.replace(".kd", "")
        if len(short) > 70:
            short = short[:67] + "..."
        print(f"| {short} | {ctr} | {n} | {mean:.6g} |")


if __name__ == "__main__":
    main()
