"""gto_hip_game_value_noalloc: the game value and its gradients with no
allocation in the call (strategy gradient staged in the library's static
device scratch) against gto_hip_game_value, which allocates its staging
buffer: bit-identical outputs on two calls in a row, the second on a
smaller spot, and a clear error for a strategy arena above the scratch's
16M entries."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _game_value(cfg, iters):
    from gtopen_amd import GameValue, GpuSolver, Spot
    from gtopen_amd.game_value import behavioural_strategy

    spot = Spot(cfg)
    g = GpuSolver(spot)
    g.iterate(iters)
    _, strat = g.download_arenas()
    g.close()
    gv = GameValue(spot)
    gv.set_strategy(*behavioural_strategy(spot, strat))
    return spot, gv


def _same(a, b):
    ua, gwa, gsa = a
    ub, gwb, gsb = b
    assert ua == ub
    for x, y in zip(gwa + gsa, gwb + gsb):
        assert (x is None) == (y is None)
        if x is not None:
            assert np.array_equal(x, y)


def test_noalloc_game_value_matches_allocating_entry(built):
    from gtopen_amd import workloads

    for name, cfg in (("turn", workloads.turn_spot()), ("river", workloads.river_spot())):
        spot, gv = _game_value(cfg, 30)
        _say(f"case: {name} spot ({spot.num_nodes} nodes, strategy arenas {spot.data_size}), both players")
        for p in (0, 1):
            ref = gv.value_and_grad(p)
            got = gv.value_and_grad(p, noalloc=True)
            _same(ref, got)
            _say(f"  player {p}: U {got[0]:.6f}, identical gradients")
        # value only, and ranges only, through the same entry
        _same(gv.value_and_grad(0, grad_weights=False, grad_strategy=False),
              gv.value_and_grad(0, grad_weights=False, grad_strategy=False, noalloc=True))
        _same(gv.value_and_grad(1, grad_strategy=False), gv.value_and_grad(1, grad_strategy=False, noalloc=True))
        gv.engine.close()


def test_noalloc_rejects_arenas_above_the_scratch(built):
    from gtopen_amd import GtoError, Spot, SpotConfig, StreetSizing, TreeConfig, workloads

    b = workloads.turn_spot()
    s = [StreetSizing.of("", ""), StreetSizing.of("25 50 75 a", "2.5x a"), StreetSizing.of("25 50 75 a", "2.5x a")]
    cfg = SpotConfig(b.board, b.range_oop, b.range_ip, TreeConfig(starting_pot=12.0, effective_stack=94.0, oop=s, ip=s))
    spot, gv = _game_value(cfg, 2)
    _say(f"case: wide turn spot, strategy arenas {spot.data_size} (> 16777216 entries)")
    assert min(spot.data_size) > 16 << 20
    with pytest.raises(GtoError, match="static device scratch"):
        gv.value_and_grad(0, noalloc=True)
    # the value and range gradients need no scratch
    ref = gv.value_and_grad(0, grad_strategy=False)
    _same(ref, gv.value_and_grad(0, grad_strategy=False, noalloc=True))
    gv.engine.close()
