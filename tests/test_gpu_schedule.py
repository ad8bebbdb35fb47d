"""Launch schedules and the autotuner. A schedule only changes how the up
sweep's terminal work is launched (runs of 16, 8 or 4 showdown nodes per
workgroup; all levels' terminals first, or each level's with that level),
so every schedule, and any sequence of schedules chosen while tuning, must
leave the arenas bit-identical to the default schedule's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANDIDATES = [(16, False), (8, False), (4, False), (16, True), (8, True), (4, True)]


def _say(msg):
    print(msg, flush=True)


def _arenas(g):
    r, s = g.download_arenas()
    return [np.asarray(a).copy() for a in (*r, *s)]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("storage", ["f32", "half"])
def test_every_schedule_gives_the_same_arenas(built, storage):
    from gtopen_amd import GpuSolver, Spot, workloads

    for name, cfg in (("river", workloads.river_spot()), ("turn", workloads.turn_spot())):
        _say(f"case: {name} spot, {storage} arenas, 20 iterations under each of 6 schedules")
        spot = Spot(cfg)
        ref = None
        for show_nodes, per_level in CANDIDATES:
            g = GpuSolver(spot, storage=storage)
            g.set_schedule(show_nodes, per_level)
            assert g.schedule() == {"show_nodes": show_nodes, "per_level": per_level}
            g.iterate(20)
            a = _arenas(g)
            e = g.exploitability()
            g.close()
            if ref is None:
                ref = (a, e)
            else:
                assert _same(ref[0], a), f"schedule ({show_nodes}, {per_level}) changed the arenas"
                assert e == ref[1]


def test_autotune_is_an_ordinary_part_of_the_solve(built):
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, autotune (1 timed iteration per candidate) vs 12 plain iterations")
    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot)
    best, ms = g.autotune(1)
    _say(f"  chose {best}; ms per iteration per candidate {['%.3f' % x for x in ms]}")
    assert (best["show_nodes"], best["per_level"]) in CANDIDATES
    assert g.schedule() == best
    assert len(ms) == 6 and all(x > 0 for x in ms)
    assert g.iteration == 12
    h = GpuSolver(spot)
    h.iterate(12)
    assert _same(_arenas(g), _arenas(h))
    g.iterate(5)
    h.iterate(5)
    assert _same(_arenas(g), _arenas(h))
    g.close()
    h.close()


def test_batch_autotune(built):
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.batch import Batch

    _say("case: river + turn spots in one batch, batch autotune vs engines iterated alone")
    spots = [Spot(workloads.river_spot()), Spot(workloads.turn_spot())]
    gs = [GpuSolver(s) for s in spots]
    b = Batch(gs)
    best, ms = b.autotune(1)
    _say(f"  chose {best}; ms per batch iteration per candidate {['%.3f' % x for x in ms]}")
    assert (best["show_nodes"], best["per_level"]) in CANDIDATES
    assert all(g.schedule() == best for g in gs)
    b.iterate(4)
    b.close()
    for s, g in zip(spots, gs):
        assert g.iteration == 16
        h = GpuSolver(s)
        h.iterate(16)
        assert _same(_arenas(g), _arenas(h))
        h.close()
        g.close()
