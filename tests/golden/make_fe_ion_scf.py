"""Writes fe_ion_scf_oracle.json: the total energy and sweep count of the
Gauss-Seidel Dirac-Fock SCF of every iron charge state of
``fe_ion_workload(seed=0)`` (Z = 26, point nucleus, ground configuration,
from the bare nucleus until the orbitals change by less than 1e-9), with
the oracle's two-electron Fock blocks and energy (oracle/scf.py).

These are the oracle's values for the GPU batched SCF test; they take CPU
minutes per ion, so they are computed once here rather than beside every GPU
run.  Run from the repository root:  python tests/golden/make_fe_ion_scf.py
"""

import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor, as_completed

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "fe_ion_scf_oracle.json")
SEED, TOL = 0, 1e-9


def _one(charge):
    sys.path.insert(0, ROOT)
    from oracle.scf import fe_ion_energy
    t0 = time.perf_counter()
    e, sweeps = fe_ion_energy(charge, seed=SEED, tol=TOL)
    return charge, e, sweeps, time.perf_counter() - t0


def main():
    for n in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        os.environ[n] = "1"
    sys.path.insert(0, ROOT)
    from atomx_amd.workloads import fe_ion_workload
    ions = fe_ion_workload(seed=SEED)
    workers = int(os.environ.get("WORKERS", min(8, os.cpu_count() or 1)))
    rows = {}
    # largest bases first so the pool ends evenly
    order = sorted(range(26), key=lambda c: -ions[c]["bsp_data"]["ns"] * len(ions[c]["kappas"]))
    with ProcessPoolExecutor(max_workers=workers) as pool:
        futs = [pool.submit(_one, c) for c in order]
        for f in as_completed(futs):
            c, e, sweeps, dt = f.result()
            bd = ions[c]["bsp_data"]
            rows[str(c)] = {"E_total_Ha": e, "sweeps": sweeps, "ns": int(bd["ns"]),
                            "ks": int(bd["ks"]), "orbitals": len(ions[c]["kappas"]),
                            "cpu_seconds": round(dt, 1)}
            print(f"Fe{c}+ ns={bd['ns']}: E = {e:.12f} Ha in {sweeps} sweeps ({dt:.0f} s)",
                  flush=True)
    doc = {"workload": "fe_ion_workload(seed=0), Z=26, point nucleus, bare-nucleus start",
           "tol": TOL, "seed": SEED,
           "ions": {k: rows[k] for k in sorted(rows, key=int)}}
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
        fh.write("\n")


if __name__ == "__main__":
    main()
