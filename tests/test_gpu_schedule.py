"""Launch schedules and the autotuner. A schedule only changes how work is
launched: the fused river sweep (one workgroup per river subtree with its
rows in LDS) or the level sweeps, and for the level sweeps the terminal
launches (runs of 16, 8 or 4 showdown nodes per workgroup; all levels'
terminals first, or each level's with that level), and whether traverser
nodes absorb their opponent children (f32 sweeps). Every schedule, and any
sequence of schedules chosen while tuning, must leave the arenas
bit-identical to the default schedule's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CANDIDATES = [(16, False, False), (8, False, False), (4, False, False), (16, True, False), (8, True, False),
              (4, True, False), (16, False, True)]
# the autotuner's candidates (absorbed opponent nodes on) and the level
# sweeps without absorbed opponent nodes
SCHEDULES = [c + (True,) for c in CANDIDATES] + [(16, False, False, False)]


def _key(sc):
    return (sc["show_nodes"], sc["per_level"], sc["river"])


def _say(msg):
    print(msg, flush=True)


def _arenas(g):
    r, s = g.download_arenas()
    return [np.asarray(a).copy() for a in (*r, *s)]


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("storage", ["f32", "half", "half_rows"])
def test_every_schedule_gives_the_same_arenas(built, storage):
    """Level sweeps under the 6 terminal schedules (and without absorbed
    opponent nodes) and the fused river sweep, 20 iterations each, on the
    river spot (one subtree: the whole tree), the turn spot (48 river
    subtrees per river chance node, raises to a cap of 10) and a report flop
    (1176 hands: six hands per thread)."""
    from gtopen_amd import GpuSolver, Spot, workloads

    flop = workloads.batched_flops(n=184)[1]
    for name, cfg in (("river", workloads.river_spot()), ("turn", workloads.turn_spot()), ("report flop", flop)):
        spot = Spot(cfg)
        _say(f"case: {name} spot ({spot.num_nodes} nodes, hands {spot.num_hands}), {storage} arenas, "
             f"20 iterations under each of {len(SCHEDULES)} schedules")
        ref = None
        for show_nodes, per_level, river, absorb in SCHEDULES:
            g = GpuSolver(spot, storage=storage)
            g.set_schedule(show_nodes, per_level, river, absorb)
            assert g.schedule() == {"show_nodes": show_nodes, "per_level": per_level, "river": river, "absorb": absorb}
            if river and storage != "half_rows":  # f16 rows: level sweeps only
                assert g.river_entries > 0
            g.iterate(20)
            a = _arenas(g)
            e = g.exploitability()
            g.close()
            if ref is None:
                ref = (a, e)
            else:
                assert _same(ref[0], a), f"schedule ({show_nodes}, {per_level}, {river}, {absorb}) changed the arenas"
                assert e == ref[1]


def test_switching_between_fused_and_level_sweeps(built):
    """The fused river sweep leaves river-level reach rows behind; a level
    sweep after it recomputes them, so switching schedules mid-solve (as the
    autotuner does) keeps the arenas bit-identical to a solve that never
    switches; locks turn the fused sweep off and unlocking turns it back on."""
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.turn_spot())
    _say("case: turn spot, fused / level / fused sweeps in one solve vs level sweeps only")
    a = GpuSolver(spot)
    b = GpuSolver(spot)
    b.set_schedule(16, False, False)
    for river in (True, False, True, False):
        a.set_schedule(16, False, river)
        a.iterate(7)
        b.iterate(7)
        assert a.exploitability() == b.exploitability()
    assert _same(_arenas(a), _arenas(b))
    a.update_locks({})
    a.iterate(3)
    b.iterate(3)
    assert _same(_arenas(a), _arenas(b))
    a.close()
    b.close()


def test_autotune_is_an_ordinary_part_of_the_solve(built):
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, autotune (1 timed iteration per candidate) vs the same number of plain iterations")
    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot)
    best, ms = g.autotune(1)
    _say(f"  chose {best}; ms per iteration per candidate {['%.3f' % x for x in ms]}")
    assert _key(best) in CANDIDATES and best["absorb"]
    assert g.schedule() == best
    assert len(ms) == len(CANDIDATES) and all(x > 0 for x in ms)
    assert g.iteration == 2 * len(CANDIDATES)
    h = GpuSolver(spot)
    h.iterate(2 * len(CANDIDATES))
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
    assert _key(best) in CANDIDATES and best["absorb"]
    assert all(g.schedule() == best for g in gs)
    b.iterate(4)
    b.close()
    n = 2 * len(CANDIDATES) + 4
    for s, g in zip(spots, gs):
        assert g.iteration == n
        h = GpuSolver(s)
        h.iterate(n)
        assert _same(_arenas(g), _arenas(h))
        h.close()
        g.close()
