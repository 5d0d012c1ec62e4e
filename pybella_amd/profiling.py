"""Summaries of rocprofv3 output databases (rocpd SQLite, ROCm 7).

``kernel_stats(db)`` returns {kernel name: (count, mean us, min us)} from a
``--kernel-trace`` run; ``pmc_means(db, counter)`` returns {kernel name: mean
counter value per dispatch} from a ``--pmc`` run.  ``python -m
pybella_amd.profiling DB [DB ...]`` prints both.
"""
from __future__ import annotations

import sqlite3
import statistics
import sys


def _rows(db, table):
    con = sqlite3.connect(db)
    cols = [r[1] for r in con.execute(f"pragma table_info({table})")]
    return cols, con.execute(f"select * from {table}").fetchall()


def kernel_stats(db):
    cols, rows = _rows(db, "kernels")
    ci = {c: i for i, c in enumerate(cols)}
    name_col = "kernel_name" if "kernel_name" in ci else "name"
    out = {}
    for r in rows:
        out.setdefault(r[ci[name_col]], []).append((r[ci["end"]] - r[ci["start"]]) / 1000.0)
    return {k: (len(v), statistics.mean(v), min(v)) for k, v in out.items()}


def pmc_means(db, counter=None):
    cols, rows = _rows(db, "pmc_events")
    out = {}
    for r in rows:
        # columns end with (..., counter name, value)
        if counter is None or r[-2] == counter:
            out.setdefault((r[5], r[-2]), []).append(r[-1])
    return {k: statistics.mean(v) for k, v in out.items()}


def main(argv):
    for db in argv:
        print(f"# {db}")
        try:
            for name, (n, mean, mn) in sorted(kernel_stats(db).items(), key=lambda kv: -kv[1][0] * kv[1][1]):
                print(f"  {n:6d} x {mean:10.3f} us (min {mn:9.3f})  {name[:120]}")
        except sqlite3.Error as e:
            print("  no kernel table:", e)
        try:
            for (name, ctr), v in pmc_means(db).items():
                print(f"  {ctr} {v:14.2f} KiB/dispatch  {name[:100]}")
        except sqlite3.Error:
            pass


if __name__ == "__main__":
    main(sys.argv[1:])
