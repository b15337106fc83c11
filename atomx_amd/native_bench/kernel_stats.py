"""Per-kernel statistics from a rocprofv3 database (the ``*_results.db``
that ``rocprofv3 --kernel-trace`` writes): calls, total and average
duration per kernel name, largest total first, as a Markdown table.

    python -m atomx_amd.native_bench.kernel_stats RESULTS.db [--match TEXT] [--top N]
"""

from __future__ import annotations

import argparse
import re
import sqlite3


def kernel_stats(db_path, match=None):
    con = sqlite3.connect(db_path)
    rows = con.execute("select name, duration from kernels").fetchall()
    con.close()
    stats = {}
    for name, dur in rows:
        short = name.replace("(anonymous namespace)::", "")
        short = re.sub(r"^void ", "", short)
        short = re.sub(r"\(.*", "", short)
        if match and match not in short:
            continue
        n, tot = stats.get(short, (0, 0.0))
        stats[short] = (n + 1, tot + float(dur))
    return sorted(((k, n, tot) for k, (n, tot) in stats.items()), key=lambda r: -r[2])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--match", default=None)
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args()
    rows = kernel_stats(a.db, a.match)
    total = sum(r[2] for r in rows)
    print("| kernel | calls | total us | avg us | share |")
    print("|---|---:|---:|---:|---:|")
    for name, n, tot in rows[:a.top]:
        print(f"| `{name}` | {n} | {tot / 1e3:.1f} | {tot / n / 1e3:.2f} | {100 * tot / total:.1f} % |")


if __name__ == "__main__":
    main()
