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


def test_python_constants_match_the_header():
    """The storage, algorithm and showdown codes the Python layer passes
    are the header's #defines."""
    from gtopen_amd import _lib

    text = open(os.path.join(ROOT, "include", "gtopen_amd.h")).read()
    defs = {k: int(v) for k, v in re.findall(r"#define\s+GTO_(\w+)\s+(-?\d+)\b", text)}
    for name, code in _lib.STORAGE.items():
        assert defs["STORAGE_" + name.upper()] == code, name
    assert (defs["ALGO_DCFR"], defs["ALGO_CFRPLUS"], defs["ALGO_PCFRPLUS"]) == (
        _lib.ALGO_DCFR, _lib.ALGO_CFRPLUS, _lib.ALGO_PCFRPLUS)
    for name, code in _lib.SHOWDOWN.items():
        assert defs["SHOWDOWN_" + name.upper()] == code, name


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


def test_out_of_device_memory_is_its_own_error(built, monkeypatch):
    """A device allocation failure reports GTO_ERR_MEMORY (not the generic
    device error), so a caller can fall back to the CPU engine as the
    reference's VRAM gate does (main.rs:654-662). GTO_INJECT_OOM=1 makes
    engine creation report one before the device is touched."""
    import ctypes as C

    from gtopen_amd import Spot, _lib as L, workloads

    spot = Spot(workloads.river_spot())
    monkeypatch.setenv("GTO_INJECT_OOM", "1")
    h = C.c_void_p()
    opts = L.HipOptions(L.STORAGE_F32, L.ALGO_DCFR, 1, 0, 1)
    rc = L.lib().gto_hip_create(spot.view_ptr, C.byref(opts), C.byref(h))
    assert rc == -3  # GTO_ERR_MEMORY
    assert b"out of device memory" in L.lib().gto_last_error()
    assert not h.value


def test_run_rejects_null_arguments(built):
    """gto_hip_run (Solver::run, cfr.rs:194-226) validates its arguments
    before touching the device."""
    import ctypes as C

    from gtopen_amd import _lib as L

    opts = L.RunOptions(10, 0.3, 5)
    rc = L.lib().gto_hip_run(None, C.byref(opts), None, L.PROGRESS_FN(0), None, None)
    assert rc == -1
    assert b"null" in L.lib().gto_last_error()
