"""Per-kernel statistics from a rocprofv3 kernel trace (the ``*_results.db``
database, or the ``*_kernel_trace.csv`` of ``--output-format csv``): calls,
total and average duration per kernel name, largest total first, as a
Markdown table.

    python -m atomx_amd.native_bench.kernel_stats TRACE [--match TEXT] [--top N]
                                                  [--last N] [--per N] [--between K]
                                                  [--until K] [--after K]

--last N keeps only the last N dispatches of each kernel (a timed loop that
follows a tuning sweep of the same kernels); --per N adds a column with the
total divided by N (time per SCF sweep, per request, ...); --between K keeps
the dispatches from the first to the last dispatch of kernel K; --until K
keeps those before the first dispatch of K, --after K those from it on
(phases of one traced run).
"""

from __future__ import annotations

import argparse
import csv
import re
import sqlite3


def _short(name):
    s = name.replace("(anonymous namespace)::", "")
    s = re.sub(r"^void ", "", s)
    return re.sub(r"\(.*", "", s)


def dispatches(path):
    """[(start ns, short kernel name, duration ns)] of every kernel dispatch."""
    if path.endswith(".csv"):
        out = []
        for r in csv.DictReader(open(path)):
            t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            out.append((t0, _short(r["Kernel_Name"]), t1 - t0))
        return out
    con = sqlite3.connect(path)
    rows = con.execute("select start, name, duration from kernels").fetchall()
    con.close()
    return [(int(t0), _short(name), int(d)) for t0, name, d in rows]


def _marks(disp, kernel):
    marks = [t0 for t0, name, _ in disp if name == kernel]
    if not marks:
        raise SystemExit(f"no dispatch of {kernel}")
    return marks


def kernel_stats(path, match=None, last=0, between=None, until=None, after=None):
    disp = sorted(dispatches(path))
    if between:  # the window from the first to the last dispatch of that kernel
        marks = _marks(disp, between)
        disp = [x for x in disp if marks[0] <= x[0] <= marks[-1]]
    if until:
        first = _marks(disp, until)[0]
        disp = [x for x in disp if x[0] < first]
    if after:
        first = _marks(disp, after)[0]
        disp = [x for x in disp if x[0] >= first]
    by_name = {}
    for t0, name, dur in disp:
        if match and match not in name:
            continue
        by_name.setdefault(name, []).append(dur)
    rows = []
    for name, durs in by_name.items():
        if last:
            durs = durs[-last:]
        rows.append((name, len(durs), float(sum(durs))))
    return sorted(rows, key=lambda r: -r[2])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--match", default=None)
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--last", type=int, default=0)
    ap.add_argument("--per", type=float, default=0.0)
    ap.add_argument("--between", default=None,
                    help="only dispatches from the first to the last one of this kernel")
    ap.add_argument("--until", default=None,
                    help="only dispatches before the first one of this kernel")
    ap.add_argument("--after", default=None,
                    help="only dispatches from the first one of this kernel on")
    a = ap.parse_args()
    rows = kernel_stats(a.trace, a.match, a.last, a.between, a.until, a.after)
    total = sum(r[2] for r in rows)
    per = f" | us per 1/{a.per:g}" if a.per else ""
    print(f"| kernel | calls | total us | avg us | share{per} |")
    print("|---|---:|---:|---:|---:|" + ("---:|" if a.per else ""))
    for name, n, tot in rows[:a.top]:
        extra = f" | {tot / 1e3 / a.per:.1f}" if a.per else ""
        print(f"| `{name}` | {n} | {tot / 1e3:.1f} | {tot / n / 1e3:.2f} | "
              f"{100 * tot / total:.1f} %{extra} |")
    print(f"\ntotal kernel time {total / 1e6:.3f} ms over {sum(r[1] for r in rows)} dispatches")


if __name__ == "__main__":
    main()
