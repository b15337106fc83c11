"""Device time per pass from a rocprofv3 CSV trace (``--kernel-trace
--memory-copy-trace --output-format csv``): the trace is cut at every
dispatch of the marker kernel (``--marker``, default ``FillFunctor``, the
fill ``serve_passes`` runs before each pass), and for each pass the table
gives its span (first start to last end of its kernels and copies), the
time at least one kernel ran, at least one host-to-device or
device-to-host copy ran, anything ran, and the span's idle rest.

    python -m atomx_amd.native_bench.pass_busy KERNEL_TRACE.csv MEMORY_COPY_TRACE.csv
                                               [--marker TEXT]
"""

from __future__ import annotations

import argparse
import csv


def _union(iv):
    """Total length of the union of [a, b) intervals."""
    tot, end = 0, None
    for a, b in sorted(iv):
        if end is None or a > end:
            tot += b - a
            end = b
        elif b > end:
            tot += b - end
            end = b
    return tot


def passes(kernel_csv, copy_csv, marker="FillFunctor"):
    """[{span, kernels, h2d, d2h, any, n_kernels, n_copies}] in ns, per pass."""
    kern, marks = [], []
    for r in csv.DictReader(open(kernel_csv)):
        a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        if marker in r["Kernel_Name"]:
            marks.append(a)
        else:
            kern.append((a, b))
    copies = []
    for r in csv.DictReader(open(copy_csv)):
        d = "h2d" if r["Direction"].endswith("HOST_TO_DEVICE") else (
            "d2h" if r["Direction"].endswith("DEVICE_TO_HOST") else "other")
        copies.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), d))
    marks.sort()
    out = []
    for i, m in enumerate(marks):
        nxt = marks[i + 1] if i + 1 < len(marks) else float("inf")
        k = [(a, b) for a, b in kern if m <= a < nxt]
        c = [(a, b, d) for a, b, d in copies if m <= a < nxt]
        every = k + [(a, b) for a, b, _ in c]
        if not every:
            continue
        span = max(b for _, b in every) - min(a for a, _ in every)
        out.append({
            "span": span, "kernels": _union(k),
            "h2d": _union([(a, b) for a, b, d in c if d == "h2d"]),
            "d2h": _union([(a, b) for a, b, d in c if d == "d2h"]),
            "any": _union(every), "n_kernels": len(k), "n_copies": len(c)})
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("kernel_csv")
    ap.add_argument("copy_csv")
    ap.add_argument("--marker", default="FillFunctor")
    args = ap.parse_args(argv)
    rows = passes(args.kernel_csv, args.copy_csv, args.marker)
    print("| pass | span ms | kernels busy ms | H2D busy ms | D2H busy ms | any busy ms | "
          "idle ms | dispatches | copies |")
    print("|---:|---:|---:|---:|---:|---:|---:|---:|---:|")
    for i, r in enumerate(rows):
        print(f"| {i} | {r['span'] / 1e6:.2f} | {r['kernels'] / 1e6:.2f} | {r['h2d'] / 1e6:.2f} | "
              f"{r['d2h'] / 1e6:.2f} | {r['any'] / 1e6:.2f} | {(r['span'] - r['any']) / 1e6:.2f} | "
              f"{r['n_kernels']} | {r['n_copies']} |")


if __name__ == "__main__":
    main()
