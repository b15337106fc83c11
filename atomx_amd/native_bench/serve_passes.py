"""Saturated passes of the serving stream's all-R^k requests through one
server, for traces: every Slater request of the 300-request Poisson stream
(seed 0) submitted at once, ``--passes`` times, the first pass cold.

    python -m atomx_amd.native_bench.serve_passes [--server native|python]
                                                   [--passes N] [--requests N]
                                                   [--max-batch N] [--max-inflight N]

Before each pass the device is synchronised and a one-element fill runs on
the current stream (a ``FillFunctor`` dispatch in a kernel trace), so
``pass_busy`` can cut the trace into passes.  Prints each pass's wall time,
the bytes moved per pass, and for the native server the scheduler thread's
host time per pass.
"""

from __future__ import annotations

import argparse
import gc
import time


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--server", choices=("native", "python"), default="native")
    ap.add_argument("--passes", type=int, default=6)
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--max-batch", type=int, default=16, help="plans per launch set")
    ap.add_argument("--max-inflight", type=int, default=4, help="launch sets in flight")
    args = ap.parse_args(argv)

    import numpy as np

    from atomx_amd import device as dv
    from atomx_amd.workloads import count_rk, serving_stream

    torch = dv.torch()
    stream = serving_stream(args.requests, rate=args.rate, seed=args.seed)
    slater = [r for _, r in stream if r.kind == "slater"]
    n_rk = sum(count_rk(r.ls, r.kmax) for r in slater)
    in_bytes = sum(8 * np.size(r.P) * (1 if r.Q is None else 2) for r in slater)
    print(f"serve_passes: {len(slater)} requests, {n_rk} R^k, {in_bytes / 1e6:.1f} MB of "
          f"orbitals in, {8 * n_rk / 1e6:.1f} MB of R^k out per pass", flush=True)
    marker = torch.zeros(1, dtype=torch.float64, device=dv.device())
    gc.collect()
    gc.freeze()
    if args.server == "native":
        from atomx_amd.serve_native import NativeSlaterServer
        srv = NativeSlaterServer(max_inflight=args.max_inflight, max_batch=args.max_batch)
    else:
        from atomx_amd.serve import SlaterServer
        srv = SlaterServer(max_inflight=args.max_inflight, max_batch=args.max_batch)
    keys = ("setup_ms", "pack_wait_ms", "alloc_ms", "queue_ms")
    try:
        prev = srv.stats() if args.server == "native" else None
        for i in range(args.passes):
            res = None
            torch.cuda.synchronize()
            marker.fill_(float(i))
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if args.server == "native":
                res = [srv.result(t) for t in [srv.submit(r) for r in slater]]
            else:
                res = [f.result() for f in [srv.submit(r) for r in slater]]
            dt = time.perf_counter() - t0
            line = f"pass {i}: {1e3 * dt:.2f} ms, {n_rk / dt:.3e} R^k/s"
            if args.server == "native":
                st = srv.stats()
                line += ", host ms " + " ".join(f"{k[:-3]} {st[k] - prev[k]:.2f}" for k in keys)
                prev = st
            print(line, flush=True)
        del res
    finally:
        srv.close()
        gc.unfreeze()
    print("serve_passes: done", flush=True)


if __name__ == "__main__":
    main()
