"""Host logic of the contraction schedules: the job tables of every order
(static K chunks and the balanced per-XCD packing) checked for coverage,
slot order and work-group ranges by a small C++ program built against the
library's headers and linked to the built library (no GPU needed)."""

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "atomx_amd", "_lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_job_tables_cover_every_tile_once(tmp_path):
    if not os.path.exists(os.path.join(LIB, "libatomx_slater.so")):
        pytest.skip("library not built")
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "jobs_check")
    build = subprocess.run(
        [HIPCC, "-O1", "-std=c++17", "--offload-arch=gfx950",
         "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "atomx_amd", "csrc"),
         os.path.join(ROOT, "tests", "jobs_check.cpp"), "-L", LIB, "-latomx_slater",
         f"-Wl,-rpath,{LIB}", "-o", exe], capture_output=True, text=True)
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(run.stdout[-2000:])
    assert run.returncode == 0 and "bad 0" in run.stdout
