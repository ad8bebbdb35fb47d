"""Batched engines (every engine's iteration graph on its own stream) and the report
runner on the GPU: batched spots must evolve exactly as when each is solved
alone, and match the CPU oracle."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _flops():
    from gtopen_amd import SpotConfig, StreetSizing, TreeConfig

    s = StreetSizing.of
    out = []
    for board, sizes in (("Ks7h2d", "33"), ("9c8c3h", "75 a"), ("AhQd5s", "50")):
        tree = TreeConfig(starting_pot=10.0, effective_stack=60.0, oop=[s(sizes, "2.5x")] * 3,
                          ip=[s(sizes, "2.5x")] * 3, max_raises=1)
        out.append(SpotConfig(board, "AA,KK,QQ,JJ,AKs,AQs,KQs,T9s,76s", "TT,99,88,AJs,KJs,QJs,JTs,98s,87s", tree))
    return out


def _arenas_equal(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[0] + a[1], b[0] + b[1]))


def test_batch_matches_solo_engines(built):
    from gtopen_amd import GpuSolver, Spot
    from gtopen_amd.batch import Batch

    spots = [Spot(c) for c in _flops()]
    _say("case: 3 flop spots (" + ", ".join(str(s.num_nodes) for s in spots) + " nodes), batch vs solo")
    solo = [GpuSolver(s) for s in spots]
    bat = [GpuSolver(s) for s in spots]
    b = Batch(bat)
    for s in solo:
        s.iterate(30)
    b.iterate(30)
    assert all(g.iteration == 30 for g in bat)
    for x, y in zip(solo, bat):
        assert _arenas_equal(x.download_arenas(), y.download_arenas())
    # an exploitability pass between batch runs leaves engines off the steady
    # state; the batch must take them through it and stay identical
    for x, y in zip(solo, bat):
        assert x.exploitability() == y.exploitability()
    for s in solo:
        s.iterate(15)
    b.iterate(15)
    for x, y in zip(solo, bat):
        assert x.iteration == y.iteration == 45
        assert _arenas_equal(x.download_arenas(), y.download_arenas())


def test_batch_matches_oracle(built):
    import oracle
    from gtopen_amd import GpuSolver, Spot
    from gtopen_amd.batch import Batch

    _say("case: batched flop spots vs the CPU oracle, 20 iterations")
    spots = [Spot(c) for c in _flops()]
    gs = [GpuSolver(s) for s in spots]
    Batch(gs).iterate(20)
    for s, g in zip(spots, gs):
        o = oracle.Oracle(s)
        o.iterate(20)
        eg, ec = g.exploitability(), o.exploitability()
        _say(f"  {s.num_nodes} nodes: gpu {eg:.6f} oracle {ec:.6f}")
        assert abs(eg - ec) <= 1e-4 * s.starting_pot


def test_report_runner(built, tmp_path):
    from gtopen_amd.reports import run_report

    _say("case: report over 3 flops, target 1% pot, cap 200 iterations")
    path = str(tmp_path / "r.json")
    rep = run_report([(c, 1) for c in _flops()], target=1.0, max_iterations=200, path=path)
    rows = rep["flops"]
    assert len(rows) == 3 and rep["flops_per_hour"] > 0
    on_disk = json.load(open(path))
    assert on_disk["complete"] and len(on_disk["flops"]) == 3
    for r in rows:
        _say(f"  {r['board']}: {r['iterations']} iterations, {r['exploit_pct']:.3f}% pot, root {r['root']}")
        assert r["exploit_pct"] <= 1.0 or r["iterations"] == 200
        assert abs(sum(r["root"]["freqs"]) - 1.0) < 1e-4
