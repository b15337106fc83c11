"""Links the native benchmark driver rk_tensor_bench (hardware-counter runs
of the tensor plan with no Python in the profiled process) against the
built library, next to it (rpath $ORIGIN).  Separate from the library build:
a failure here never fails the library.

    python -m atomx_amd.native_bench.build
"""

from __future__ import annotations

import os
import subprocess

from .._build import HIPCC, LIB_DIR, ROOT

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE_BENCH = os.path.join(LIB_DIR, "rk_tensor_bench")


def build() -> str:
    src = os.path.join(HERE, "rk_tensor_bench.cpp")
    tmp = NATIVE_BENCH + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), src, "-L", LIB_DIR,
           "-latomx_slater", "-Wl,-rpath,$ORIGIN", "-o", tmp]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError("hipcc failed on the native benchmark:\n" + proc.stdout + proc.stderr)
    os.replace(tmp, NATIVE_BENCH)
    return NATIVE_BENCH


if __name__ == "__main__":
    print(build())
