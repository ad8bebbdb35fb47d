"""f16 rows (GTO_STORAGE_HALF_ROWS): the DCFR iterations keep reach and value
rows in f16 with a power-of-two scale per row, over the f16 arenas of
GTO_STORAGE_HALF. Every row entry is rounded once to f16 (2^-11 relative),
so the engine follows the f32-row engine closely but not bit for bit; the
FP64 oracle judges its trajectory. Sweeps other than the DCFR iteration
(best response, node values, the game value) keep f32 rows, and running
them between iterations must not change the solve."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CURRENT = 2  # node_values mode: current strategies


def _say(msg):
    print(msg, flush=True)


def _arenas(g):
    r, s = g.download_arenas()
    return [np.asarray(a, np.float64).copy() for a in (*r, *s)]


def _odd_turn_spot():
    """A turn spot with odd hand counts (301 / 457 combos): the update then
    takes the hands one at a time (with an even count it reads and writes
    them in pairs)."""
    from gtopen_amd import Spot, SpotConfig, StreetSizing, TreeConfig, card_str, workloads as W

    rng = np.random.default_rng(5)
    board = W._board(rng, 4)
    free = W._free_combos(board)
    rngs = []
    for n in (301, 457):
        rngs.append(W._range_text(sorted(int(x) for x in rng.choice(free, size=n, replace=False)), W._beta(rng, n)))
    s = [StreetSizing.of("", ""), StreetSizing.of("33 75 a", "2.5x a"), StreetSizing.of("33 75 a", "2.5x a")]
    tree = TreeConfig(starting_pot=12.0, effective_stack=94.0, oop=s, ip=[StreetSizing(**vars(x)) for x in s])
    spot = Spot(SpotConfig("".join(card_str(c) for c in board), rngs[0], rngs[1], tree))
    assert spot.num_hands[0] % 2 == 1 and spot.num_hands[1] % 2 == 1
    return spot


def _spots():
    from gtopen_amd import Spot, workloads

    return (("river", Spot(workloads.river_spot())), ("turn", Spot(workloads.turn_spot())),
            ("odd-hand turn", _odd_turn_spot()))


def test_half_rows_follow_f32_rows(built):
    """One iteration from zero arenas on the river and turn spots: the f16-row
    engine's decoded arenas against the f32-row engine's (same f16 arena
    format), each within 1% of the arena's largest entry (measured: about
    1e-3; every row entry carries one f16 rounding per tree level). A wrong
    row scale, a stale row or a row read in the wrong format moves entries
    by the full scale. Later iterations are compared by exploitability
    only: regret matching turns near-zero regrets (actions a hand is
    indifferent between) into different strategies under any perturbation,
    so single entries part ways after the first iteration."""
    from gtopen_amd import GpuSolver

    for name, spot in _spots():
        _say(f"case: {name} spot {tuple(spot.num_hands)}, 1 iteration, f16 rows vs f32 rows (f16 arenas)")
        a, b = GpuSolver(spot, storage="half"), GpuSolver(spot, storage="half_rows")
        a.iterate(1)
        b.iterate(1)
        for k, (x, y) in enumerate(zip(_arenas(a), _arenas(b))):
            scale = np.abs(x).max()
            err = np.abs(x - y).max()
            _say(f"  arena {k}: max |f16 rows - f32 rows| {err:.3e} (scale {scale:.3e})")
            assert scale > 0
            assert err <= 1e-2 * scale
        a.close()
        b.close()


def test_half_rows_trajectory_against_fp64(built):
    """River spot for 200 iterations and turn spot for 100, exploitability
    every 25 against the FP64 oracle's trajectory, within 1e-3 of the pot
    (the band of the BF16 showdown form, whose reach operands carry 8 bits;
    f16 rows carry 11). The odd-hand turn spot is left to the one-iteration
    check: its trajectory is chaotic enough that the CPU f32 engine itself
    strays 0.14% of the pot from FP64 by iteration 75 (f16 rows: at most
    0.11% over 200 iterations, f16 arenas with f32 rows 0.18%)."""
    import oracle
    from gtopen_amd import GpuSolver

    for (name, spot), n in zip(_spots()[:2], (200, 100)):
        _say(f"case: {name} spot {tuple(spot.num_hands)}, f16 rows vs FP64 oracle, {n} iterations, check every 25")
        g, ref = GpuSolver(spot, storage="half_rows"), oracle.Oracle(spot, precision=64)
        pot = spot.starting_pot
        for k in range(n // 25):
            g.iterate(25)
            ref.iterate(25)
            eg, e64 = g.exploitability(), ref.exploitability()
            _say(f"  iter {25 * (k + 1)}: f16 rows {eg:.6f} fp64 {e64:.6f} ({abs(eg - e64) / pot * 100:.4f}% pot)")
            assert abs(eg - e64) <= 1e-3 * pot
        g.close()


def test_f32_row_sweeps_between_iterations_leave_the_solve_unchanged(built):
    """Node values of the current strategies (a down sweep of the same
    strategies, in f32 rows), root values and exploitability between
    iterations: the f16-row engine recomputes its rows afterwards, so its
    arenas stay bit-identical to an engine that only iterates."""
    from gtopen_amd import GpuSolver, Spot, _lib as L, workloads

    spot = Spot(workloads.turn_spot())
    _say("case: turn spot, f16 rows, 12 iterations with f32-row sweeps between them vs 12 plain iterations")
    show = np.nonzero(spot.kind == L.KIND_SHOWDOWN)[0][:16]
    a, b = GpuSolver(spot, storage="half_rows"), GpuSolver(spot, storage="half_rows")
    a.iterate(12)
    for k in range(12):
        b.iterate(1)
        b.node_values(k % 2, CURRENT, show)
        b.root_cfv(1 - k % 2, CURRENT)
        if k % 4 == 3:
            b.exploitability()
    assert a.iteration == b.iteration == 12
    assert all(np.array_equal(x, y) for x, y in zip(_arenas(a), _arenas(b)))
    assert a.exploitability() == b.exploitability()


def test_half_rows_batch_matches_solo(built):
    """Two report flops through one batch graph, f16 rows: bit-identical to the
    same engines iterated alone."""
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.batch import Batch

    spots = [Spot(c) for c in workloads.batched_flops(n=184)[:2]]
    _say("case: 2 report flops, f16 rows, batch vs solo, 12 iterations")
    solo = [GpuSolver(s, storage="half_rows") for s in spots]
    bat = [GpuSolver(s, storage="half_rows") for s in spots]
    for g in solo:
        g.iterate(12)
    Batch(bat).iterate(12)
    for x, y in zip(solo, bat):
        assert all(np.array_equal(u, v) for u, v in zip(_arenas(x), _arenas(y)))


def test_half_rows_locked_node(built):
    """A locked node (lock weights bound the children's rows): the f16-row
    engine's locked solve tracks the f32-row engine's."""
    from gtopen_amd import GpuSolver, Spot, SpotConfig, StreetSizing, TreeConfig

    s = StreetSizing.of
    tree = TreeConfig(starting_pot=100.0, effective_stack=100.0, oop=[s("", ""), s("", ""), s("100", "")],
                      ip=[s("", ""), s("", ""), s("", "")])
    spot = Spot(SpotConfig("QcJc9c3d2d", "AcKc,8h7h", "QdQh", tree))
    bet = [{"type": "action", "index": 1}]
    _say("case: clairvoyance river, IP locked to always call, 300 iterations, f16 rows vs f32 rows")
    out = []
    for st in ("half", "half_rows"):
        g = GpuSolver(spot, storage=st)
        g.iterate(50)
        g.lock_node(bet, {"kind": "range", "freqs": [0.0, 1.0]}, "ip always calls")
        g.iterate(300)
        out.append((g.average_strategy(0), g.exploitability()))
        g.close()
    _say(f"  exploitability: f32 rows {out[0][1]:.6f}, f16 rows {out[1][1]:.6f}")
    np.testing.assert_allclose(out[1][0], out[0][0], atol=2e-3)
    assert abs(out[1][1] - out[0][1]) <= 1e-3 * spot.starting_pot


def test_half_rows_reject_matrix_core_showdown(built):
    from gtopen_amd import GpuSolver, GtoError, Spot, workloads

    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot, storage="half_rows")
    for form in ("bf16", "fp8"):
        with pytest.raises(GtoError, match="f16 rows"):
            g.set_showdown(form)
    g.set_showdown("exact")
    g.iterate(2)
    g.close()
