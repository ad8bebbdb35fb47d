"""GPU parity: the HIP engine against the CPU oracle and the reference's
golden trajectory. Exploitability must agree within 1e-4 of the pot
(f32 arithmetic on both sides; the GPU sums some f64 totals in a different
order, so per-element arenas agree to float rounding, not bit for bit;
test_gpu_sequential.py covers the mode that sums in the CPU's order and
matches it bit for bit).
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

POT_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)


def _close_arenas(a, b, scale, rtol=2e-4):
    for p in range(2):
        x, y = a[p], b[p]
        err = np.abs(x.astype(np.float64) - y.astype(np.float64))
        lim = rtol * np.abs(y.astype(np.float64)) + rtol * scale
        assert np.all(err <= lim), f"player {p}: max err {err.max()} at {int(err.argmax())}"


def test_river_trajectory_matches_oracle(built):
    """The river configuration at its full length: 1000 iterations,
    exploitability logged every 50, within 1e-4 of the pot of the CPU
    oracle's at every check; arenas after 100 iterations entry by entry, and
    after 1000 bit for bit (on this spot no f64 terminal sum rounds
    differently in the GPU's association than in the CPU's sequential one,
    so the two engines stay identical; measured)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: river spot (400x400), 1000 iterations, check every 50")
    spot = Spot(workloads.river_spot())
    gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
    pot = spot.starting_pot
    gpu.iterate(100)
    cpu.iterate(100)
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _close_arenas(rg, rc, pot)
    _close_arenas(sg, sc, 1.0)
    worst = 0.0
    for it in range(150, 1001, 50):
        gpu.iterate(it - gpu.iteration)
        cpu.iterate(it - cpu.iteration)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        worst = max(worst, abs(eg - ec) / pot)
        _say(f"  iter {it}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) / pot < POT_TOL
    _say(f"  largest difference over 1000 iterations: {worst:.2e} of the pot")
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    for p in range(2):
        assert np.array_equal(rg[p], rc[p]) and np.array_equal(sg[p], sc[p]), f"player {p}: arenas differ at 1000"


def test_turn_matches_oracle(built):
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot (1128x1128, 28509 nodes), 12 iterations")
    spot = Spot(workloads.turn_spot())
    gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
    pot = spot.starting_pot
    for n in (2, 10):
        gpu.iterate(n)
        cpu.iterate(n)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) / pot < POT_TOL
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _close_arenas(rg, rc, pot)
    _close_arenas(sg, sc, 1.0)


# bench/new_f32_dcfr.log:2-4 of the reference: the CPU f32 engine on
# bench_spot.json (Ks7h2d, 1,245,971 nodes), exploitability in chips.
GOLDEN_F32 = {10: 33.7904, 20: 9.3381, 30: 4.8733}


def test_bench_spot_reproduces_reference_trajectory(built):
    from gtopen_amd import GpuSolver, Spot

    _say("case: bench_spot.json (reference golden trajectory), 30 iterations")
    spot = Spot(open(golden("bench_spot.json")).read())
    gpu = GpuSolver(spot)
    for it in (10, 20, 30):
        gpu.iterate(it - gpu.iteration)
        e = gpu.exploitability()
        _say(f"  iter {it}: gpu {e:.4f} chips, reference {GOLDEN_F32[it]:.4f}")
        assert abs(e - GOLDEN_F32[it]) / spot.starting_pot < POT_TOL


def test_best_response_matches_oracle_on_gpu_arenas(built):
    """GPU best response and average-strategy values on the same arenas as the
    CPU (best_response.rs:19-223), per hand."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, BR/avg root values on GPU-produced arenas")
    spot = Spot(workloads.turn_spot())
    gpu = GpuSolver(spot)
    gpu.iterate(6)
    r, s = gpu.download_arenas()
    cpu = oracle.Oracle(spot)
    cpu.set_arenas(r, s)
    for p in range(2):
        for mode in (0, 1):
            vg = gpu.root_cfv(p, mode).astype(np.float64)
            vc = cpu.root_cfv(p, mode)
            err = np.abs(vg - vc).max()
            _say(f"  p{p} mode{mode}: max |gpu-cpu| {err:.3e}")
            assert err <= 1e-5 * spot.starting_pot + 1e-5 * np.abs(vc).max()
    assert abs(gpu.exploitability() - cpu.exploitability()) / spot.starting_pot < 1e-6


def test_graph_replay_matches_eager(built):
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: river spot, hipGraph replay vs eager launches, 40 iterations with checks between")
    spot = Spot(workloads.river_spot(seed=11))
    a = GpuSolver(spot, use_graphs=True)
    b = GpuSolver(spot, use_graphs=False)
    for _ in range(4):
        a.iterate(10)
        b.iterate(10)
        assert a.exploitability() == b.exploitability()
    ra, sa = a.download_arenas()
    rb, sb = b.download_arenas()
    for p in range(2):
        assert np.array_equal(ra[p], rb[p]) and np.array_equal(sa[p], sb[p])


# bench/new_i16_dcfr.log:2-3 of the reference: the CPU compressed (i16/u16)
# engine on bench_spot.json.
GOLDEN_I16 = {10: 33.5076, 20: 9.2809}


def test_compressed_river_matches_oracle(built):
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: river spot, compressed i16/u16 arenas, 60 iterations")
    spot = Spot(workloads.river_spot())
    gpu, cpu = GpuSolver(spot, storage="compressed"), oracle.Oracle(spot, storage="compressed")
    pot = spot.starting_pot
    for _ in range(6):
        gpu.iterate(10)
        cpu.iterate(10)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) / pot < POT_TOL
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _close_arenas(rg, rc, pot, rtol=2e-3)
    _close_arenas(sg, sc, 1.0, rtol=2e-3)


def test_compressed_bench_spot_reproduces_reference_trajectory(built):
    from gtopen_amd import GpuSolver, Spot

    _say("case: bench_spot.json, compressed arenas, 20 iterations")
    spot = Spot(open(golden("bench_spot.json")).read())
    gpu = GpuSolver(spot, storage="compressed")
    for it in (10, 20):
        gpu.iterate(it - gpu.iteration)
        e = gpu.exploitability()
        _say(f"  iter {it}: gpu {e:.4f} chips, reference {GOLDEN_I16[it]:.4f}")
        assert abs(e - GOLDEN_I16[it]) / spot.starting_pot < POT_TOL


def test_full_range_river_matches_oracle(built):
    """All 1081 combos per player on a river board (the largest showdown
    lists: the showdown kernel runs with fewer wavefronts per workgroup)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, SpotConfig, StreetSizing, TreeConfig

    _say("case: river, full 1081x1081 ranges, 30 iterations")
    s = [StreetSizing(), StreetSizing(), StreetSizing.of("50 a", "2.5x")]
    tree = TreeConfig(starting_pot=10.0, effective_stack=40.0, oop=s, ip=s, max_raises=1)
    spot = Spot(SpotConfig("Kd8s5h2c2d", "22+,A2+,K2+,Q2+,J2+,T2+,92+,82+,72+,62+,52+,42+,32",
                           "22+,A2+,K2+,Q2+,J2+,T2+,92+,82+,72+,62+,52+,42+,32", tree))
    assert spot.num_hands == (1081, 1081)
    gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
    for _ in range(3):
        gpu.iterate(10)
        cpu.iterate(10)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) / spot.starting_pot < POT_TOL
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _close_arenas(rg, rc, spot.starting_pot)
    _close_arenas(sg, sc, 1.0)


def test_long_card_lists_with_few_hands_match_oracle(built):
    """Ace-heavy ranges on a river board: few hands per player (172 / 316)
    but up to 43 opponent hands holding one card, longer than the showdown
    kernel's short card-list form takes (36), so the long form runs with
    the few-hands row layout."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, SpotConfig, StreetSizing, TreeConfig

    _say("case: river, A2+ vs A2+,Q2+ (long card lists, few hands), 30 iterations")
    s = [StreetSizing(), StreetSizing(), StreetSizing.of("50 a", "2.5x")]
    tree = TreeConfig(starting_pot=10.0, effective_stack=40.0, oop=s, ip=s, max_raises=1)
    spot = Spot(SpotConfig("Kd8s5h2c2d", "A2+", "A2+,Q2+", tree))
    assert max(spot.num_hands) <= 768
    longest = 0
    for q in range(2):
        cnt = np.zeros(52, dtype=np.int64)
        np.add.at(cnt, np.asarray(spot.hand_c1[q], dtype=np.int64), 1)
        np.add.at(cnt, np.asarray(spot.hand_c2[q], dtype=np.int64), 1)
        longest = max(longest, int(cnt.max()))
    assert longest > 36
    gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
    for _ in range(3):
        gpu.iterate(10)
        cpu.iterate(10)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) / spot.starting_pot < POT_TOL
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _close_arenas(rg, rc, spot.starting_pot)
    _close_arenas(sg, sc, 1.0)


def test_cfr_plus_matches_oracle(built):
    """CFR+ (cfr.rs:57-61: positive regrets kept, negative ones floored by a
    zero discount, strategy weights (t/(t+1))^2) on the GPU engine."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, CFR+, 40 iterations, check every 10")
    spot = Spot(workloads.turn_spot())
    gpu, cpu = GpuSolver(spot, algo="cfr+"), oracle.Oracle(spot, algo="cfr+")
    pot = spot.starting_pot
    for k in range(4):
        gpu.iterate(10)
        cpu.iterate(10)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {10 * (k + 1)}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) <= POT_TOL * pot


def test_wide_action_menus_match_oracle(built):
    """Nodes with 5 to 8 actions (the update kernels' generic path): a river
    spot with six bet sizes and two raise sizes, f32 and f16 arenas, against
    the oracle (f16: against FP64, within the f16 engine's tolerance)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, SpotConfig, StreetSizing, TreeConfig, workloads, _lib as L

    base = workloads.river_spot()
    s = [StreetSizing(), StreetSizing(), StreetSizing.of("25 33 50 75 125 a", "2.5x 4x")]
    cfg = SpotConfig(base.board, base.range_oop, base.range_ip,
                     TreeConfig(starting_pot=20.0, effective_stack=90.0, oop=s, ip=s, max_raises=1))
    spot = Spot(cfg)
    na = np.asarray(spot.num_children)[np.asarray(spot.kind) == L.KIND_ACTION]
    _say(f"case: river spot, action counts {np.bincount(na).tolist()}, 40 iterations")
    assert na.max() >= 7
    gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
    pot = spot.starting_pot
    for k in range(4):
        gpu.iterate(10)
        cpu.iterate(10)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {10 * (k + 1)}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) <= POT_TOL * pot
    gpu.close()
    gh, o64 = GpuSolver(spot, storage="half"), oracle.Oracle(spot, precision=64)
    gh.iterate(40)
    o64.iterate(40)
    eh, e64 = gh.exploitability(), o64.exploitability()
    _say(f"  f16 arenas, iter 40: gpu {eh:.6f} fp64 {e64:.6f}")
    assert abs(eh - e64) <= 1e-3 * pot
    gh.close()


def test_pcfr_plus_matches_oracle(built):
    """PCFR+ (cfr.rs:250-289, 337-348: CFR+ discounts, the current strategy
    matched on max(R,0) plus the last instantaneous regret) on the GPU
    engine, which stores that sum per entry, against the oracle; river and
    turn spots, exploitability every 10 iterations, and the arenas after
    two iterations, bit for bit on the river spot."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: river spot, PCFR+, arenas after 2 iterations")
    spot = Spot(workloads.river_spot())
    gpu, cpu = GpuSolver(spot, algo="pcfr+"), oracle.Oracle(spot, algo="pcfr+")
    gpu.iterate(2)
    cpu.iterate(2)
    (gr, gs), (cr, cs) = gpu.download_arenas(), cpu.arenas()
    for g, c in zip((*gr, *gs), (*cr, *cs)):
        scale = max(1.0, float(np.abs(c).max()))
        assert np.allclose(g, c, rtol=0, atol=1e-5 * scale)
    gpu.close()
    for name, cfg, n in (("river", workloads.river_spot(), 60), ("turn", workloads.turn_spot(), 40)):
        _say(f"case: {name} spot, PCFR+, {n} iterations, check every 10")
        spot = Spot(cfg)
        gpu, cpu = GpuSolver(spot, algo="pcfr+"), oracle.Oracle(spot, algo="pcfr+")
        pot = spot.starting_pot
        for k in range(n // 10):
            gpu.iterate(10)
            cpu.iterate(10)
            eg, ec = gpu.exploitability(), cpu.exploitability()
            _say(f"  iter {10 * (k + 1)}: gpu {eg:.6f} cpu {ec:.6f}")
            assert abs(eg - ec) <= POT_TOL * pot
        gpu.close()


def test_pcfr_plus_needs_f32_arenas(built):
    from gtopen_amd import GpuSolver, GtoError, Spot, workloads

    spot = Spot(workloads.river_spot())
    for st in ("compressed", "half"):
        with pytest.raises(GtoError, match="PCFR"):
            GpuSolver(spot, storage=st, algo="pcfr+")


@pytest.mark.timeout(900)
def test_turn_long_trajectory_compressed(built):
    """The turn configuration's full length with the compressed engine: 500
    iterations on the full-range turn spot against the oracle every 50:
    quantization snaps both engines to the same values, so the trajectories
    agree to the tolerance."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, compressed arenas, 500 iterations, check every 50")
    spot = Spot(workloads.turn_spot())
    gpu, cpu = GpuSolver(spot, storage="compressed"), oracle.Oracle(spot, storage="compressed")
    pot = spot.starting_pot
    for k in range(10):
        gpu.iterate(50)
        cpu.iterate(50)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {50 * (k + 1)}: gpu {eg:.6f} cpu {ec:.6f}")
        assert abs(eg - ec) <= POT_TOL * pot


@pytest.mark.timeout(900)
def test_turn_long_trajectory_f32_against_fp64(built):
    """The turn configuration's full length: 500 f32 iterations on the
    full-range turn spot, checked every 50. Over long runs two f32 engines
    drift apart at the level of f32 rounding (the CPU f32 engine itself sits
    ~2e-3 chips from the FP64 trajectory here), so the GPU is held to the
    FP64 oracle: it may deviate from FP64 by no more than the CPU f32 engine
    does, plus 1e-4 of the pot; the first 50 iterations must also agree with
    the CPU f32 engine within 1e-4 of the pot."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, f32 and f16 arenas, 500 iterations vs CPU f32 and FP64 oracles, check every 50")
    spot = Spot(workloads.turn_spot())
    gpu, cpu, ref = GpuSolver(spot), oracle.Oracle(spot), oracle.Oracle(spot, precision=64)
    half = GpuSolver(spot, storage="half")
    pot = spot.starting_pot
    for k in range(10):
        for e in (gpu, cpu, ref, half):
            e.iterate(50)
        eg, ec, e64, eh = gpu.exploitability(), cpu.exploitability(), ref.exploitability(), half.exploitability()
        _say(f"  iter {50 * (k + 1)}: gpu {eg:.6f} cpu f32 {ec:.6f} fp64 {e64:.6f} gpu f16 {eh:.6f} "
             f"(gpu-cpu {abs(eg - ec) / pot:.2e} of the pot)")
        assert abs(eg - e64) <= abs(ec - e64) + POT_TOL * pot
        # the f16 arenas (an engine format of this build) are held to the
        # same band around the FP64 trajectory
        assert abs(eh - e64) <= abs(ec - e64) + POT_TOL * pot
        if k < 1:
            assert abs(eg - ec) <= POT_TOL * pot


def test_half_arena_transfer(built, tmp_path):
    """f16 arenas: uploads round to f16 against a per-node power-of-two
    scale (relative error <= 2^-11 for every value above 2^-14 of its node's
    largest), downloads decode exactly, and a save/load round trip through
    the f32 file keeps the exploitability."""
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.save import load_solver

    _say("case: river spot, f16 arenas: upload/download and save/load")
    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot, storage="half")
    g.iterate(40)
    e0 = g.exploitability()
    r, s = g.download_arenas()
    g.upload_arenas(r, s)  # values already on the f16 grid: unchanged
    r2, s2 = g.download_arenas()
    for x, y in zip(r + s, r2 + s2):
        assert np.array_equal(x, y)
    assert g.exploitability() == e0
    rng = np.random.default_rng(9)
    x = [rng.standard_normal(spot.data_size[q]).astype(np.float32) for q in range(2)]
    g.upload_arenas(x, None)
    y, _ = g.download_arenas()
    for q in range(2):
        big = np.abs(x[q]) > 1e-3
        rel = np.abs(y[q][big] - x[q][big]) / np.abs(x[q][big])
        assert rel.max() <= 2.0 ** -11
    g2 = GpuSolver(spot, storage="half")
    g2.iterate(40)
    path = str(tmp_path / "h.gto")
    g2.save(path)
    _, g3, _ = load_solver(path, storage="half")
    assert abs(g3.exploitability() - g2.exploitability()) < 1e-9
