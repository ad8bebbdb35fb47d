"""The C-ABI library loads and exports every function include/gtopen_amd.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT


def _declared():
    text = open(os.path.join(ROOT, "include", "gtopen_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gto_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(built):
    from gtopen_amd import _lib

    names = _declared()
    assert len(names) >= 25
    h = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(h, n)]
    assert not missing, missing
    bound = {n for n, _, _ in _lib.SIGNATURES}
    assert set(names) <= bound, set(names) - bound


def test_engine_reports_device_errors_without_crashing(built):
    """Without a GPU the engine's create fails with an error string, never a
    silent CPU path."""
    import pytest

    from gtopen_amd import GpuSolver, GtoError, Spot, workloads

    try:
        import torch

        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    spot = Spot(workloads.river_spot())
    with pytest.raises(GtoError):
        GpuSolver(spot)
