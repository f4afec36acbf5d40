"""Kernel-time summary of a rocprofv3 result: a rocpd SQLite database or the
kernel_stats.csv of a --kernel-trace --stats --output-format csv run.

Usage: python3 rocpd_summary.py <results.db | kernel_stats.csv> [top]
Prints, per kernel name: calls, total ms, average us, min/max us, share.
"""
import csv
import sqlite3
import sys


def _rows(path):
    if path.endswith(".csv"):
        with open(path) as f:
            rows = [(r["Name"], int(r["Calls"]), float(r["TotalDurationNs"]), float(r["AverageNs"]),
                     float(r["MinNs"]), float(r["MaxNs"])) for r in csv.DictReader(f)]
        return sorted(rows, key=lambda r: -r[2])
    c = sqlite3.connect(path)
    return c.execute(
        "select name, count(*), sum(duration), avg(duration), min(duration), max(duration) "
        "from kernels group by name order by sum(duration) desc").fetchall()


def summary(db, top=25):
    rows = _rows(db)
    total = sum(r[2] for r in rows) or 1
    out = [f"{'kernel':<60} {'calls':>6} {'total_ms':>10} {'avg_us':>10} {'min_us':>10} "
           f"{'max_us':>10} {'pct':>6}"]
    for name, n, tot, avg, mn, mx in rows[:top]:
        short = name if len(name) <= 60 else name[:57] + "..."
        out.append(f"{short:<60} {n:>6} {tot / 1e6:>10.3f} {avg / 1e3:>10.2f} {mn / 1e3:>10.2f} "
                   f"{mx / 1e3:>10.2f} {100.0 * tot / total:>6.2f}")
    return "\n".join(out)


if __name__ == "__main__":
    print(summary(sys.argv[1], int(sys.argv[2]) if len(sys.argv) > 2 else 25))
