"""Summarise a rocprofv3 ``--kernel-trace`` database into a per-kernel table.

    python -m nrpy_amd.profsummary gpurun_out/prof/run_results.db
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


if __name__ == "__main__":
    main()
