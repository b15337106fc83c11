"""The C-ABI library loads and exports every symbol its headers declare (CPU)."""

import glob
import os
import re

import pytest

from atomx_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(ax_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "ax_yk_batch" in names and "ax_rk_tensor_run" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, f"symbols declared but not exported: {missing}"


def test_python_binding_covers_every_declared_symbol():
    _lib.load()
    assert set(declared_symbols()) <= set(_lib.SIGNATURES)


def test_version_and_host_tables():
    import numpy as np

    import oracle as O
    lib = _lib.load()
    assert lib.ax_version() >= 100
    from atomx_amd.grid import partial_matrices
    for ng in (5, 7, 8, 9, 16):
        mo, mi = partial_matrices(ng)
        omo, omi = O.yk_partial_matrices(ng)
        assert np.max(np.abs(mo - omo)) < 2e-15
        assert np.max(np.abs(mi - omi)) < 2e-15


def test_errors_map_to_reference_exceptions():
    import ctypes

    import numpy as np
    mo = np.empty((1,))
    with pytest.raises(ValueError, match="n_gauss"):
        _lib.call("ax_yk_partial_matrices", 0, mo.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                  mo.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def test_no_cpu_fallback_without_gpu():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from atomx_amd import yk
    import numpy as np
    with pytest.raises(RuntimeError, match="GPU"):
        yk.compute_yk(np.ones((4, 8)), np.ones((4, 8)), np.ones((4, 8)), 4, 8, 0)
