"""CPU-side checks of the drop-in boundary: libicc loads and exports every
entry point include/icc.h declares (no compute call needs a GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "icc.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(icc_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ["icc_apply", "icc_apply_dev", "icc_pcg", "icc_diag", "icc_energies", "icc_sens", "icc_solve_ir",
              "icc_filter_sens", "icc_filter_fwd", "icc_filter_bwd", "icc_wq_diag", "icc_classify_cylinders"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    import immersed_catmull_clark_amd as m
    L = ctypes.CDLL(m.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    from immersed_catmull_clark_amd._abi import SIGNATURES
    bound = {s[0] for s in SIGNATURES}
    missing = [s for s in declared_symbols() if s not in bound]
    assert not missing, missing


def test_library_reports_errors_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present: covered by the gpu tests")
    import immersed_catmull_clark_amd as m
    with pytest.raises(m.IccError):
        m.Context(0)
