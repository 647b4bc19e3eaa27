"""The C-ABI library: loads, and exports every entry point declared in
include/lbm_d2q9.h (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from lbm_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "lbm_d2q9.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "lbm_step" in names and "lbm_macro" in names and len(names) >= 30


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(_lib.exported_names()) == _declared()


def test_abi_version():
    assert _lib.load().lbm_abi_version() == 1


def test_no_gpu_means_no_kernels():
    """The product path fails loudly without a GPU: there is no CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from lbm_amd.lattice import lattice
    from tests.cases import small_cavity
    lat = lattice(small_cavity())
    with pytest.raises(RuntimeError):
        lat.macro()
