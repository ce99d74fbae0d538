"""The C ABI library loads and exports every entry point include/pdx.h declares
(no compute call: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    inc = os.path.join(ROOT, "include")
    text = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pdx_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("pdx_create", "pdx_destroy", "pdx_set_particles", "pdx_build_bonds", "pdx_shape_tensor",
                 "pdx_compute_F_stress", "pdx_compute_forces", "pdx_dr_solve", "pdx_break_bonds",
                 "pdx_commit_state", "pdx_get", "pdx_set", "pdx_last_error", "pdx_create_batch"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from simul_amd import _native
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    # the Python binding covers the same surface
    assert set(declared()) <= set(_native.SIGNATURES) | {"pdx_last_error"}


def test_no_gpu_fails_loudly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from simul_amd import _native
    with pytest.raises(RuntimeError):
        _native.Context(8, 3)
