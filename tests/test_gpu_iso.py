"""Suit isomorphism on the GPU engine (plan.rs:154-191, kernels.cu:272-300,
cfr.rs:568-660): one chance branch per suit orbit is traversed, chance
nodes read the other cards' values from the representative's row through
the symmetry's hand table, and downloads and strategies of the other
branches are the representatives' suit-permuted copies (cfr.rs:749-870).
Checked against the oracle's isomorphism, which is itself checked against
its every-branch traversal."""
import numpy as np
import pytest

from test_oracle_iso import symmetric_turn

pytestmark = pytest.mark.gpu

POT_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)


def _flat(arenas):
    return [*arenas[0], *arenas[1]]


@pytest.mark.parametrize("storage", ["f32", "compressed"])
def test_iso_engine_matches_oracle(built, storage):
    import oracle
    from gtopen_amd import GpuSolver, Spot

    _say(f"case: symmetric turn spot (6 suit symmetries), {storage} arenas, iso GPU vs iso oracle")
    spot = Spot(symmetric_turn())
    g, o = GpuSolver(spot, storage=storage), oracle.Oracle(spot, storage=storage)
    g.iterate(2)
    o.iterate(2)
    for x, y in zip(_flat(g.download_arenas()), _flat(o.arenas())):
        scale = max(1.0, float(np.abs(y).max()))
        err = float(np.abs(x - y).max())
        assert err <= 1e-5 * scale, (err, scale)
    pot = spot.starting_pot
    for k in range(4):
        g.iterate(10)
        o.iterate(10)
        eg, eo = g.exploitability(), o.exploitability()
        _say(f"  iter {12 + 10 * k}: gpu {eg:.6f} oracle {eo:.6f}")
        assert abs(eg - eo) <= POT_TOL * pot
    g.close()


def test_iso_against_every_branch_on_the_gpu(built):
    from gtopen_amd import GpuSolver, Spot

    _say("case: symmetric turn spot, GPU with and without isomorphism")
    spot = Spot(symmetric_turn())
    a, b = GpuSolver(spot), GpuSolver(spot, isomorphism=False)
    _say(f"  arena entries traversed: iso {a.reached_entries}, every branch {b.reached_entries}")
    assert a.reached_entries < b.reached_entries
    # the same arenas while regret matching has no near-zero sign to flip;
    # afterwards the two stay within f32 rounding of the same trajectory
    a.iterate(2)
    b.iterate(2)
    for x, y in zip(_flat(a.download_arenas()), _flat(b.download_arenas())):
        scale = max(1.0, float(np.abs(y).max()))
        assert float(np.abs(x - y).max()) <= 1e-5 * scale
    pot = spot.starting_pot
    for k in range(4):
        a.iterate(10)
        b.iterate(10)
        assert abs(a.exploitability() - b.exploitability()) <= POT_TOL * pot
    a.close()
    b.close()


def test_strategy_of_a_copied_branch(built):
    import oracle
    from gtopen_amd import GpuSolver, Spot, _lib as L

    _say("case: symmetric turn spot, solved strategy of a non-representative node")
    spot = Spot(symmetric_turn())
    g, o = GpuSolver(spot), oracle.Oracle(spot)
    g.iterate(15)
    o.iterate(15)
    root_chance = next(n for n in range(spot.num_nodes) if spot.kind[n] == L.KIND_CHANCE)
    cs = int(spot.children_start[root_chance])
    three_s = 1 * 4 + 3  # card = 4 * rank + suit, suits c d h s: its representative is 3c
    dst = int(spot.children[cs + three_s])
    # the node and one below it
    for n in (dst, spot.child(dst, 0)):
        if spot.kind[n] != L.KIND_ACTION:
            continue
        gs = g.solved_strategy(n)
        os_ = o.average_strategy(n)
        assert np.allclose(gs, os_, rtol=0, atol=1e-5)
        # the downloaded (symmetrized) strategy sums give the same strategy
        p = int(spot.player[n])
        nh, na = spot.num_hands[p], int(spot.num_children[n])
        s = g.download_arenas()[1][p][int(spot.data_offset[n]):int(spot.data_offset[n]) + na * nh]
        s = np.maximum(s.reshape(na, nh).astype(np.float64), 0.0)
        tot = s.sum(axis=0)
        ref = np.where(tot > 1e-12, s / np.where(tot > 0, tot, 1.0), 1.0 / na).astype(np.float32).reshape(-1)
        assert np.allclose(gs, ref, rtol=0, atol=1e-6)
    g.close()
