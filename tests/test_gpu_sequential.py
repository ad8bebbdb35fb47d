"""The sequential-sums showdown mode: the fold and showdown f64 sums in the
CPU engine's order wherever the order could change a bit, so the GPU's
arenas and exploitability equal the CPU oracle's bit for bit over whole
configurations, not just to a tolerance. Each check compares every arena
entry with np.array_equal and the exploitability with ==."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _identical(gpu, cpu, what, values=False):
    """Arenas and exploitability equal; values: also the root rows of the
    best response (mode 0) and of the average strategy (mode 1), f32 on
    both sides."""
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    for p in range(2):
        for name, a, b in (("regrets", rg[p], rc[p]), ("strategy sums", sg[p], sc[p])):
            d = int(np.count_nonzero(a != b))
            assert d == 0, f"{what}: player {p} {name}: {d} of {a.size} entries differ"
    eg, ec = gpu.exploitability(), cpu.exploitability()
    _say(f"  {what}: arenas identical, exploitability gpu {eg!r} cpu {ec!r}")
    assert eg == ec
    if values:
        for p in range(2):
            for mode in (0, 1):
                vg, vc = gpu.root_cfv(p, mode), cpu.root_cfv(p, mode).astype(np.float32)
                assert np.array_equal(vg, vc), f"{what}: root values of player {p}, mode {mode}"
        _say(f"  {what}: root best-response and average-strategy rows identical")


@pytest.mark.timeout(900)
def test_turn_full_length_bit_identical(built):
    """The turn configuration at its full length (500 iterations), f32 and
    i16/u16 arenas: the GPU stays identical to the oracle at every check
    (with the default f64 association the f32 arenas part after about 10
    iterations)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.turn_spot())
    for storage, marks in (("f32", (10, 100, 300, 500)), ("compressed", (10, 100, 200))):
        _say(f"case: turn spot, {storage} arenas, sequential sums, {marks[-1]} iterations")
        gpu = GpuSolver(spot, storage=storage, showdown="sequential")
        cpu = oracle.Oracle(spot, storage=storage)
        for k in marks:
            gpu.iterate(k - gpu.iteration)
            cpu.iterate(k - cpu.iteration)
            if storage == "f32":
                _identical(gpu, cpu, f"iter {k}", values=k == marks[-1])
            else:
                eg, ec = gpu.exploitability(), cpu.exploitability()
                _say(f"  iter {k}: exploitability gpu {eg!r} cpu {ec!r}")
                assert eg == ec
        gpu.close()


@pytest.mark.timeout(900)
def test_flop_srp_bit_identical(built):
    """The headline flop SRP spot (1.45M nodes, 2352 river boards): 20
    iterations identical to the oracle (the default association parts after
    the first)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: flop SRP spot, f32 arenas, sequential sums, 20 iterations")
    spot = Spot(workloads.flop_srp_spot())
    gpu, cpu = GpuSolver(spot, showdown="sequential"), oracle.Oracle(spot)
    for k in (2, 10, 20):
        gpu.iterate(k - gpu.iteration)
        cpu.iterate(k - cpu.iteration)
        _identical(gpu, cpu, f"iter {k}", values=k == 20)
    gpu.close()


def test_algorithms_rake_and_reference_spot_bit_identical(built):
    """CFR+, PCFR+ and raked DCFR on the turn spot (50 iterations) and the
    reference's bench_spot flop (30 iterations), identical to the oracle."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, SpotConfig, TreeConfig, workloads

    base = workloads.turn_spot()
    t = base.tree
    raked = SpotConfig(base.board, base.range_oop, base.range_ip,
                       TreeConfig(starting_pot=t.starting_pot, effective_stack=t.effective_stack, oop=t.oop, ip=t.ip,
                                  rake_pct=0.05, rake_cap=3.0))
    cases = (("turn, CFR+", Spot(base), "cfr+", 50), ("turn, PCFR+", Spot(base), "pcfr+", 50),
             ("turn, 5% rake capped at 3", Spot(raked), "dcfr", 50),
             ("bench_spot", Spot(open(golden("bench_spot.json")).read()), "dcfr", 30))
    for name, spot, algo, n in cases:
        _say(f"case: {name}, sequential sums, {n} iterations")
        gpu = GpuSolver(spot, algo=algo, showdown="sequential")
        cpu = oracle.Oracle(spot, algo=algo)
        gpu.iterate(n)
        cpu.iterate(n)
        _identical(gpu, cpu, f"iter {n}", values=True)
        gpu.close()


def test_locked_solve_bit_identical(built):
    """A node lock (query.rs:702-776, the root raked to a target mix) on the
    turn spot: 20 iterations, the lock on both engines, 30 more, identical
    to the oracle."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, root locked to a mix after 20 iterations, sequential sums, 50 iterations")
    spot = Spot(workloads.turn_spot())
    gpu, cpu = GpuSolver(spot, showdown="sequential"), oracle.Oracle(spot)
    gpu.iterate(20)
    cpu.iterate(20)
    na = int(spot.num_children[0])
    freqs = [float(k + 1) for k in range(na)]
    gpu.lock_node([], {"kind": "range", "freqs": freqs}, "root mix")
    cpu.set_locks(gpu.locks)
    gpu.iterate(30)
    cpu.iterate(30)
    _identical(gpu, cpu, "iter 50", values=True)
    gpu.close()


def test_sequential_needs_f32_rows(built):
    """The sequential sums are f64 sums of f32 reach rows: refused with f16
    rows; switching back to the default association is allowed."""
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot, storage="half_rows")
    with pytest.raises(Exception):
        g.set_showdown("sequential")
    g.close()
    g = GpuSolver(spot, showdown="sequential")
    g.iterate(3)
    g.set_showdown("exact")
    g.iterate(3)
    assert g.iteration == 6
    g.close()
