"""The solve-cli driver's output lines parse with the reference's chart
script regex (bench/make_chart.py:34-36 of the reference), restated here
because the line format is the contract between the two programs."""
import re

from gtopen_amd.cli import done_line, header_line, progress_line
from gtopen_amd.solver import Progress

# bench/make_chart.py:34-36
ITER_RE = r"iter\s+(\d+)\s+exploitability\s+([\d.]+) chips \(([\d.]+)% pot\)\s+\[([\d.]+)s(?:, gpu ([\d.]+)s)?\]"


def test_progress_lines_parse_with_the_chart_regex():
    for p in (Progress(10, 33.79041, 56.317, 1.234, 0.91), Progress(12345, 0.0123, 0.0205, 1234.5, 1000.25)):
        line = progress_line(p)
        m = re.match(ITER_RE, line)
        assert m, line
        assert int(m.group(1)) == p.iteration
        assert abs(float(m.group(2)) - p.exploit_chips) < 1e-4
        assert abs(float(m.group(3)) - p.exploit_pct_pot) < 1e-3
        assert abs(float(m.group(4)) - p.elapsed_secs) < 0.05
        assert abs(float(m.group(5)) - p.gpu_secs) < 0.05


def test_header_and_done_lines():
    h = header_line(1.234, 1245971, 500748, (257, 319), 2494.8e6, "F32", "Dcfr")
    assert h == ("tree built in 1.23s: 1245971 nodes (500748 action nodes), hands 257/319, arenas 2494.8 MB, "
                 "storage F32, algo Dcfr")
    # the chart script reads the arena size from the header (make_chart.py:45)
    assert re.search(r"arenas ([\d.]+) MB", h).group(1) == "2494.8"
    d = done_line(Progress(320, 0.0165, 0.2998, 3.75, 3.6), 2048.0)
    assert d == "done: iter 320 exploitability 0.300% pot in 3.8s (gpu kernels 3.60s = 11 ms/iter)  peak_rss 2048 MB"
