"""Summarises the counter runs of counters.sh: kernel durations, calibrated
HBM bytes and GB/s per kernel, MFMA work and busy fraction of the
contraction, and the traffic table the benchmark reads.

    python -m atomx_amd.native_bench.summarize OUTDIR [--table TABLE]

Byte counters are scaled by the factor that makes the run's closing
256 MiB device copy read and write 256 MiB.  FETCH_SIZE and WRITE_SIZE are
KiB per dispatch.  SQ_INSTS_VALU_MFMA_MOPS_* count MFMA work in units of 512
operations (one v_mfma_f64_16x16x4_f64 is 4 units); SQ_VALU_MFMA_BUSY_CYCLES
is summed over the SIMDs, so its busy fraction is taken against 1024 SIMDs x
the kernel's traced duration x 2.4 GHz (the published peak engine clock).
"""

from __future__ import annotations

import argparse
import collections
import csv
import os
import re

import numpy as np

from ..workloads import valid_pairs

HBM_PEAK = 8.0e12
CLOCK = 2.4e9
SIMDS = 1024
COPY_BYTES = 256 << 20
KERNELS = ("pair_density_batch_kernel", "yk_scan", "rk_gemm", "rk_reduce_kernel")


def short(name: str) -> str:
    name = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name)


def tiles_of(path: str) -> int:
    with open(path, "rb") as f:
        nv, ng, n_orb, kmax = np.fromfile(f, dtype=np.int32, count=4)
        f.seek(16 + 8 * (2 * nv * ng + 2 * n_orb * nv * ng))
        ls = np.fromfile(f, dtype=np.int32, count=n_orb)
    return sum(((len(valid_pairs(ls, k)) + 63) // 64) * ((len(valid_pairs(ls, k)) + 63) // 64 + 1)
               // 2 for k in range(kmax + 1))


def per_kernel(path: str, value: str):
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        if "Grid_Size" in r:
            grid = int(r["Grid_Size"])
        else:  # kernel trace: work-items per dimension
            grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        key = (short(r["Kernel_Name"]), grid)
        if value == "duration":
            out[key]["duration"].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9)
        else:
            out[key][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def mean(v):
    return sum(v) / len(v) if v else float("nan")


def summarise(d: str, prec: str, n_tiles: int):
    trace = per_kernel(os.path.join(d, f"{prec}-trace", "nb_kernel_trace.csv"), "duration")
    fetch = per_kernel(os.path.join(d, f"{prec}-fetch", "nb_counter_collection.csv"), "counter")
    write = per_kernel(os.path.join(d, f"{prec}-write", "nb_counter_collection.csv"), "counter")
    mfma = per_kernel(os.path.join(d, f"{prec}-mfma", "nb_counter_collection.csv"), "counter")
    copy = ("__amd_rocclr_copyBuffer", COPY_BYTES // 2048)
    f_fetch = COPY_BYTES / (1024 * mean(fetch[copy]["FETCH_SIZE"]))
    f_write = COPY_BYTES / (1024 * mean(write[copy]["WRITE_SIZE"]))
    rows = []
    for key in sorted(trace):
        name, grid = key
        if not any(name.startswith(k) for k in KERNELS):
            continue
        t = mean(trace[key]["duration"])
        fb = mean(fetch[key]["FETCH_SIZE"]) * 1024 * f_fetch if key in fetch else float("nan")
        wb = mean(write[key]["WRITE_SIZE"]) * 1024 * f_write if key in write else float("nan")
        split = None
        if name.startswith("rk_gemm"):  # static K split, or the balanced order's level
            groups = grid // 256
            split = (str(groups // n_tiles) if groups % n_tiles == 0
                     else f"bal{groups // 64}")
        row = {"kernel": name, "grid": grid, "split": split, "dispatches": len(trace[key]["duration"]),
               "us": t * 1e6, "fetch": fb, "write": wb, "gbs": (fb + wb) / t / 1e9}
        m = mfma.get(key, {})
        if name.startswith("rk_gemm") and m:
            mops = mean(m.get("SQ_INSTS_VALU_MFMA_MOPS_F64", [])) + mean(
                m.get("SQ_INSTS_VALU_MFMA_MOPS_I8", []))
            busy = mean(m.get("SQ_VALU_MFMA_BUSY_CYCLES", []))
            row["ops"] = mops * 512
            row["busy"] = busy / (SIMDS * t * CLOCK)
        rows.append(row)
    return f_fetch, f_write, rows


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("outdir")
    ap.add_argument("--table", default=None, help="also write the traffic table here")
    a = ap.parse_args(argv)
    n_tiles = tiles_of(os.path.join(a.outdir, "w.bin"))
    table = []
    for prec in ("fp64", "split"):
        ff, fw, rows = summarise(a.outdir, prec, n_tiles)
        print(f"## {prec} ({n_tiles} upper 64x64 tiles; calibration: fetch x{ff:.4f}, write x{fw:.4f})")
        print()
        print("| kernel | grid | K split | dispatches | avg us | fetch MB | write MB | HBM GB/s "
              "(% of 8 TB/s) | MFMA Gop | MFMA busy |")
        print("|---|---|---|---|---|---|---|---|---|---|")
        for r in rows:
            ops = f"{r['ops'] / 1e9:.3f}" if "ops" in r else ""
            busy = f"{100 * r['busy']:.1f} %" if "busy" in r else ""
            print(f"| `{r['kernel']}` | {r['grid']} | {r['split'] or ''} | {r['dispatches']} | "
                  f"{r['us']:.2f} | {r['fetch'] / 1e6:.1f} | {r['write'] / 1e6:.1f} | "
                  f"{r['gbs']:.0f} ({100 * r['gbs'] * 1e9 / HBM_PEAK:.1f} %) | {ops} | {busy} |")
            if r["split"]:
                table.append((r["kernel"], r["split"], r["fetch"], r["write"], r["us"]))
        print()
    if a.table:
        with open(a.table, "w") as f:
            f.write("# HBM traffic per launch of the R^k contraction kernels on the benchmark tensor "
                    "(40 orbitals, 400 x 8 nodes, k = 0..8),\n# one line per K split, from "
                    "rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of the native driver "
                    "(atomx_amd/native_bench/counters.sh),\n# scaled by the closing 256 MiB copy; "
                    "avg_us from the kernel trace of the same driver.\n"
                    "# columns: kernel schedule (K split of the xcd-contiguous order, or balN: the\n"
                    "# balanced order at level N) fetch_bytes write_bytes traffic_bytes avg_us\n")
            for k, s, fb, wb, us in table:
                f.write(f"traffic: {k} {s} {fb:.0f} {wb:.0f} {fb + wb:.0f} {us:.2f}\n")


if __name__ == "__main__":
    main()
