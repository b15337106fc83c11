"""CPU: the trace summaries used for the serving profile (pass_busy) on a
hand-made trace whose per-pass busy times are known."""

import csv

from atomx_amd.native_bench import pass_busy


def _write(path, header, rows):
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def test_union_of_intervals():
    assert pass_busy._union([]) == 0
    assert pass_busy._union([(0, 10), (5, 15), (20, 25)]) == 20
    assert pass_busy._union([(5, 6), (0, 10)]) == 10  # nested, unsorted


def test_passes_cut_at_markers(tmp_path):
    k = tmp_path / "k.csv"
    c = tmp_path / "c.csv"
    kh = ["Kernel_Name", "Start_Timestamp", "End_Timestamp"]
    _write(k, kh, [
        ["void at::native::FillFunctor<double>(...)", 0, 1],      # pass 0 marker
        ["rk_gemm_f64_kernel(...)", 10, 30],
        ["yk_scan(...)", 20, 40],                                # overlaps: kernels busy 30
        ["void at::native::FillFunctor<double>(...)", 100, 101],  # pass 1 marker
        ["rk_gemm_f64_kernel(...)", 110, 120],
    ])
    ch = ["Direction", "Start_Timestamp", "End_Timestamp"]
    _write(c, ch, [
        ["MEMORY_COPY_HOST_TO_DEVICE", 5, 15],    # overlaps the first kernel
        ["MEMORY_COPY_DEVICE_TO_HOST", 40, 50],
        ["MEMORY_COPY_HOST_TO_DEVICE", 130, 135],
    ])
    rows = pass_busy.passes(str(k), str(c))
    assert len(rows) == 2
    a, b = rows
    assert (a["span"], a["kernels"], a["h2d"], a["d2h"], a["any"]) == (45, 30, 10, 10, 45)
    assert (a["n_kernels"], a["n_copies"]) == (2, 2)
    assert (b["span"], b["kernels"], b["h2d"], b["d2h"], b["any"]) == (25, 10, 5, 0, 15)
