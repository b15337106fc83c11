"""Writes the benchmark tensor's workload (the synthetic grid, orbitals and
angular momenta bench.py times) to a binary file for rk_tensor_bench:
int32 nv, ng, n_orb, kmax; float64 r, w (nv*ng); float64 P, Q (n_orb, nv*ng);
int32 l (n_orb).

    python -m atomx_amd.native_bench.write_workload OUT [--nv 400 --ng 8 --n-orb 40 --kmax 8]
"""

from __future__ import annotations

import argparse

import numpy as np

from ..workloads import synthetic_grid, synthetic_orbitals


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--nv", type=int, default=400)
    ap.add_argument("--ng", type=int, default=8)
    ap.add_argument("--n-orb", type=int, default=40)
    ap.add_argument("--kmax", type=int, default=8)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    gr, grw = synthetic_grid(nv=a.nv, ng=a.ng, r_min=1e-5, r_max=60.0)
    P, Q, ls = synthetic_orbitals(gr, grw, n_orb=a.n_orb, seed=a.seed)
    with open(a.out, "wb") as f:
        np.array([a.nv, a.ng, a.n_orb, a.kmax], dtype=np.int32).tofile(f)
        for x in (gr, grw, P, Q):
            np.ascontiguousarray(x, dtype=np.float64).tofile(f)
        np.ascontiguousarray(ls, dtype=np.int32).tofile(f)


if __name__ == "__main__":
    main()
