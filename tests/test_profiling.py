"""The rocprofv3 summarizer (gtopen_amd.profiling) on small synthetic CSVs
laid out as rocprofv3 writes them: kernel names, per-iteration figures and
the copy calibration of the HBM byte counters."""
import csv
import json
import os

import pytest

from gtopen_amd import profiling as P


def _write(path, header, rows):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w", newline="") as f:
        w = csv.writer(f)
        w.writerow(header)
        w.writerows(rows)


def _pmc(root, counter, values):
    rows = [[i, f'"{k}"', counter, v] for i, (k, v) in enumerate(values)]
    _write(os.path.join(root, "x", "run_counter_collection.csv"),
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)


UP = "void gto::k_up_action<0, 2, 3>(gto::DevTree, gto::DevState, gto::ActRec const*, unsigned int, int, int)"
SHOW = "void gto::k_up_show<3, 9, 0>(gto::DevTree, gto::DevState, gto::ShowRec const*, unsigned int, int, int)"


def test_short_names():
    assert P.short_name(UP) == "k_up_action<0, 2, 3>"
    assert P.short_name('"gto::k_begin_iteration(gto::DevState)"') == "k_begin_iteration"
    assert P.short_name("__amd_rocclr_fillBufferAligned") == "__amd_rocclr_fillBufferAligned"


def test_kernel_table_per_iteration(tmp_path):
    _write(str(tmp_path / "t" / "a" / "run_kernel_stats.csv"), ["Name", "Calls", "TotalDurationNs"],
           [[UP, 40, 8_000_000], [SHOW, 4, 2_000_000]])
    md = P.kernels_markdown(str(tmp_path / "t"), 2)
    lines = md.splitlines()
    assert lines[2] == "| `k_up_action<0, 2, 3>` | 40 | 8.00 | 4.000 | 200.0 |"
    assert lines[3] == "| `k_up_show<3, 9, 0>` | 4 | 2.00 | 1.000 | 500.0 |"


def test_traffic_scales_by_the_copy(tmp_path):
    # a 1 MiB copy reads FETCH_SIZE 512 KiB (scale 2) and writes WRITE_SIZE 1024 KiB (scale 1)
    _pmc(str(tmp_path / "cf"), "FETCH_SIZE", [("copy", 512.0)])
    _pmc(str(tmp_path / "cw"), "WRITE_SIZE", [("copy", 1024.0)])
    _pmc(str(tmp_path / "f"), "FETCH_SIZE", [(UP, 100.0), (UP, 300.0), (SHOW, 50.0)])
    _pmc(str(tmp_path / "w"), "WRITE_SIZE", [(UP, 40.0)])
    t = P.traffic(str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"), 2, 1)
    assert t["fetch_scale"] == pytest.approx(2.0) and t["write_scale"] == pytest.approx(1.0)
    up = t["kernels"]["k_up_action<0, 2, 3>"]
    assert up["read_bytes_per_iter"] == pytest.approx(400 * 1024 * 2 / 2)
    assert up["write_bytes_per_iter"] == pytest.approx(40 * 1024 / 2)
    assert t["kernels"]["k_up_show<3, 9, 0>"]["write_bytes_per_iter"] == 0.0
    assert P.main(["traffic", str(tmp_path / "f"), str(tmp_path / "w"), str(tmp_path / "cf"), str(tmp_path / "cw"),
                   "2", "1", "spot"]) == 0


def test_counters_with_conflict_share(tmp_path):
    root = str(tmp_path / "lds")
    rows = []
    for i, (c, v) in enumerate((("SQ_LDS_BANK_CONFLICT", 30.0), ("SQ_LDS_IDX_ACTIVE", 120.0), ("SQ_WAVE_CYCLES", 9.0))):
        rows.append([i, f'"{SHOW}"', c, v])
    rows.append([9, f'"{UP}"', "SQ_WAVE_CYCLES", 50.0])
    _write(os.path.join(root, "x", "run_counter_collection.csv"),
           ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"], rows)
    md = P.counters_markdown(root, 3).splitlines()
    assert md[0].endswith("| bank-conflict share |")
    assert md[2].startswith("| `k_up_action<0, 2, 3>` |") and md[2].endswith("| - |")
    assert md[3].startswith("| `k_up_show<3, 9, 0>` |") and md[3].endswith("| 0.250 |")


def test_usage_on_bad_arguments(capsys):
    assert P.main(["nonsense"]) == 2
    assert "profiling" in capsys.readouterr().err
