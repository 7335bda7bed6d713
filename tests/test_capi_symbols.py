"""The C-ABI library loads and exports every function include/fjmodels.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def declared():
    src = open(os.path.join(ROOT, "include", "fjmodels.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(fj_\w+)\s*\(", src, re.M)))


def test_header_lists_entry_points():
    names = declared()
    assert "fj_compute_new_phi" in names and "fj_moments" in names
    assert len(names) >= 20


def test_library_exports_all_declared_symbols():
    from fjmodels_amd import capi
    assert os.path.exists(capi.LIB_PATH), "libfjmodels.so not built"
    lib = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == declared()


def test_no_device_is_an_error_not_a_fallback():
    """Without a GPU the context cannot be created; nothing silently computes on the CPU."""
    import pytest
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import fjmodels_amd as fj
    with pytest.raises(fj.FJError):
        fj.Context(0)
