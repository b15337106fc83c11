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


HBM_PEAK = 8.0e12
CLOCK = 2.4e9
SIMDS = 1024
COPY_BYTES = 256 << 20
KERNELS = ("pair_density_batch_kernel", "yk_scan", "rk_gemm", "rk_reduce_kernel")


def short(name: str) -> str:
    name = name.replace("void ", "").replace("(anonymous namespace)::", "")
    return re.sub(r"\(.*", "", name)


ORDERS = ("tile-major", "xcd-contiguous", "xcd-contiguous-split-fastest",
          "dynamic-heaviest-first", "balanced-k-ranges")


def schedules_of(log: str):
    """The (order, K split) of every schedule the driver ran, in run order."""
    out = []
    for line in open(log):
        m = re.search(r"order (\d+), K split (\d+):", line)
        if line.startswith("workload") and m:
            out.append((int(m.group(1)), int(m.group(2))))
    return out


def dispatches(path: str, value: str):
    """{dispatch id: (short kernel name, grid, {counter or 'duration': value})}."""
    out = {}
    for r in csv.DictReader(open(path)):
        if "Grid_Size" in r:
            grid = int(r["Grid_Size"])
        else:  # kernel trace: work-items per dimension
            grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        did = int(r["Dispatch_Id"])
        name, _, vals = out.setdefault(did, (short(r["Kernel_Name"]), grid, {}))
        if value == "duration":
            vals["duration"] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-9
        else:
            vals[r["Counter_Name"]] = vals.get(r["Counter_Name"], 0.0) + float(r["Counter_Value"])
    return out


def per_key(disp, scheds):
    """Values per key: the contraction kernel's dispatches keyed by the
    schedule they ran under (the driver runs the same number per schedule,
    in the order of its log), every other kernel by (name, grid)."""
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    gemm = sorted(d for d, (name, _, _) in disp.items() if name.startswith("rk_gemm"))
    per = len(gemm) // len(scheds) if scheds else 0
    if per == 0 or per * len(scheds) != len(gemm):
        raise SystemExit(f"{len(gemm)} contraction dispatches do not split over "
                         f"{len(scheds)} schedules")
    sched_of = {d: scheds[i // per] for i, d in enumerate(gemm)}
    for d, (name, grid, vals) in disp.items():
        key = (name, "%d:%d" % sched_of[d]) if d in sched_of else (name, grid)
        for c, v in vals.items():
            out[key][c].append(v)
    return out


def mean(v):
    return sum(v) / len(v) if v else float("nan")


def summarise(d: str, prec: str):
    scheds = schedules_of(os.path.join(d, f"{prec}-trace.log"))
    load = lambda name, kind, v: per_key(dispatches(os.path.join(d, f"{prec}-{name}", kind), v),
                                         scheds)
    trace = load("trace", "nb_kernel_trace.csv", "duration")
    fetch = load("fetch", "nb_counter_collection.csv", "counter")
    write = load("write", "nb_counter_collection.csv", "counter")
    mfma = load("mfma", "nb_counter_collection.csv", "counter")
    extra = {}
    for name in ("lds", "stall"):  # optional passes
        path = os.path.join(d, f"{prec}-{name}", "nb_counter_collection.csv")
        if os.path.exists(path):
            for key, vals in load(name, "nb_counter_collection.csv", "counter").items():
                extra.setdefault(key, {}).update(vals)
    copy = ("__amd_rocclr_copyBuffer", COPY_BYTES // 2048)
    f_fetch = COPY_BYTES / (1024 * mean(fetch[copy]["FETCH_SIZE"]))
    f_write = COPY_BYTES / (1024 * mean(write[copy]["WRITE_SIZE"]))
    rows = []
    for key in sorted(trace, key=str):
        name, where = key
        if not any(name.startswith(k) for k in KERNELS):
            continue
        t = mean(trace[key]["duration"])
        fb = mean(fetch[key]["FETCH_SIZE"]) * 1024 * f_fetch if key in fetch else float("nan")
        wb = mean(write[key]["WRITE_SIZE"]) * 1024 * f_write if key in write else float("nan")
        sched = where if isinstance(where, str) else None
        row = {"kernel": name, "where": where, "schedule": sched,
               "dispatches": len(trace[key]["duration"]),
               "us": t * 1e6, "fetch": fb, "write": wb, "gbs": (fb + wb) / t / 1e9}
        m = mfma.get(key, {})
        if name.startswith("rk_gemm") and m:
            mops = mean(m.get("SQ_INSTS_VALU_MFMA_MOPS_F64", [])) + mean(
                m.get("SQ_INSTS_VALU_MFMA_MOPS_I8", []))
            busy = mean(m.get("SQ_VALU_MFMA_BUSY_CYCLES", []))
            row["ops"] = mops * 512
            row["busy"] = busy / (SIMDS * t * CLOCK)
        x = {c: mean(v) for c, v in extra.get(key, {}).items()}
        wc = x.get("SQ_WAVE_CYCLES")
        if wc:
            row["waves"] = {
                "lds_active_inst": x.get("SQ_ACTIVE_INST_LDS", float("nan")) / wc,
                "wait_inst_lds": x.get("SQ_WAIT_INST_LDS", float("nan")) / wc,
                "wait_any": x.get("SQ_WAIT_ANY", float("nan")) / wc,
                "wait_inst_any": x.get("SQ_WAIT_INST_ANY", float("nan")) / wc,
                "active_valu": x.get("SQ_ACTIVE_INST_VALU", float("nan")) / wc,
                "active_vmem": x.get("SQ_ACTIVE_INST_VMEM", float("nan")) / wc,
                "active_any": x.get("SQ_ACTIVE_INST_ANY", float("nan")) / wc,
                "lds_bank_conflict_per_active": x.get("SQ_LDS_BANK_CONFLICT", float("nan"))
                / max(x.get("SQ_LDS_IDX_ACTIVE", float("nan")), 1.0),
                "lds_idx_active_per_cu_cycle": x.get("SQ_LDS_IDX_ACTIVE", float("nan"))
                / (256 * t * CLOCK),
            }
        rows.append(row)
    return f_fetch, f_write, rows


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("outdir")
    ap.add_argument("--table", default=None, help="also write the traffic table here")
    ap.add_argument("--top", type=int, default=10,
                    help="contraction schedules listed per mode, fastest first (the table "
                         "holds all)")
    a = ap.parse_args(argv)
    table = []
    for prec in ("fp64", "split"):
        ff, fw, rows = summarise(a.outdir, prec)
        print(f"## {prec} (calibration: fetch x{ff:.4f}, write x{fw:.4f})")
        print()
        print("| kernel | grid, or schedule order:split | dispatches | avg us | fetch MB | "
              "write MB | HBM GB/s (% of 8 TB/s) | MFMA Gop | MFMA busy |")
        print("|---|---|---|---|---|---|---|---|---|")
        fast = sorted((r for r in rows if r["schedule"]), key=lambda r: r["us"])[:a.top]
        listed = [r for r in rows if not r["schedule"]] + fast
        for r in listed:
            ops = f"{r['ops'] / 1e9:.3f}" if "ops" in r else ""
            busy = f"{100 * r['busy']:.1f} %" if "busy" in r else ""
            where = r["where"]
            if r["schedule"]:
                o, sp = map(int, where.split(":"))
                where = f"{ORDERS[o]}:{sp}"
            print(f"| `{r['kernel']}` | {where} | {r['dispatches']} | "
                  f"{r['us']:.2f} | {r['fetch'] / 1e6:.1f} | {r['write'] / 1e6:.1f} | "
                  f"{r['gbs']:.0f} ({100 * r['gbs'] * 1e9 / HBM_PEAK:.1f} %) | {ops} | {busy} |")
        for r in rows:
            if r["schedule"]:
                table.append((r["kernel"], r["schedule"], r["fetch"], r["write"], r["us"],
                              r.get("busy", float("nan"))))
        n_sched = sum(1 for r in rows if r["schedule"])
        print(f"\n({len(fast)} fastest of {n_sched} contraction schedules listed)\n")
        waved = [r for r in listed if "waves" in r]
        if waved:
            keys = list(waved[0]["waves"])
            print("Wave cycles (fractions of SQ_WAVE_CYCLES; LDS bank-conflict cycles per "
                  "LDS-active cycle; LDS-active cycles per CU cycle):\n")
            print("| kernel | where | " + " | ".join(keys) + " |")
            print("|---|---|" + "---|" * len(keys))
            for r in waved:
                print(f"| `{r['kernel']}` | {r['where']} | "
                      + " | ".join(f"{r['waves'][k]:.3f}" for k in keys) + " |")
            print()
    if a.table:
        with open(a.table, "w") as f:
            f.write("# HBM traffic per launch of the R^k contraction kernels on the benchmark tensor "
                    "(40 orbitals, 400 x 8 nodes, k = 0..8),\n# one line per schedule the tuner "
                    "can choose, from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of the "
                    "native driver\n# (atomx_amd/native_bench/counters.sh), scaled by the closing "
                    "256 MiB copy; avg_us from the kernel trace of the same driver.\n"
                    "# columns: kernel order:split (order index into tile-major, xcd-contiguous,\n"
                    "# xcd-contiguous-split-fastest, dynamic-heaviest-first, balanced-k-ranges)\n"
                    "# fetch_bytes write_bytes traffic_bytes avg_us mfma_busy\n")
            for k, s_, fb, wb, us, busy in table:
                f.write(f"traffic: {k} {s_} {fb:.0f} {wb:.0f} {fb + wb:.0f} {us:.2f} {busy:.3f}\n")


if __name__ == "__main__":
    main()
