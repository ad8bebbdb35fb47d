"""GPU parity on the benchmark workloads at their real sizes: the flop SRP
spot (1,445,723 nodes) in f32 against the oracle and in every precision
form judged by the FP64 oracle; raked exploitability and raked DCFR
trajectories (the reference's crates/solver/tests/gpu.rs:133-165); batched
report flops through the batch path; run-to-target and the solve-cli
driver's output.
"""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

POT_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)


def _close_arenas_chunked(a, b, scale, rtol=2e-4, chunk=1 << 24):
    """Per-entry agreement |x - y| <= rtol |y| + rtol scale, in chunks (the
    flop arenas hold 390M entries per player)."""
    for p in range(2):
        x, y = a[p], b[p]
        worst = 0.0
        for k in range(0, x.size, chunk):
            xs = x[k:k + chunk].astype(np.float64)
            ys = y[k:k + chunk].astype(np.float64)
            err = np.abs(xs - ys)
            lim = rtol * np.abs(ys) + rtol * scale
            bad = np.nonzero(err > lim)[0]
            assert bad.size == 0, f"player {p}: entry {k + int(bad[0])}: {xs[bad[0]]} vs {ys[bad[0]]}"
            worst = max(worst, float((err / lim).max()) if err.size else 0.0)
        _say(f"  player {p}: worst |gpu-cpu| / tolerance {worst:.3f}")


def _rel_l2(a, b, chunk=1 << 24):
    """||a - b|| / ||b|| over a whole arena, in chunks."""
    num = den = 0.0
    for k in range(0, a.size, chunk):
        x = a[k:k + chunk].astype(np.float64)
        y = b[k:k + chunk].astype(np.float64)
        num += float(np.dot(x - y, x - y))
        den += float(np.dot(y, y))
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


@pytest.mark.timeout(900)
def test_flop_srp_f32_matches_oracle(built):
    """The headline workload: f32 iterations on the GPU and on the CPU
    oracle. After 2 iterations the arenas agree entry by entry (the only
    difference is the association of the f64 terminal sums); exploitability
    within 1e-4 of the pot at 10 and 20. Over 20 iterations of a 1.4M-node
    tree f32 rounding differences are amplified where regret matching
    switches, so two f32 engines drift apart (here ~3e-3 relative L2 in the
    regrets, with exploitability still within 4e-6 of the pot): at 20 the
    GPU's arenas must be as close to the FP64 oracle's as the CPU f32
    engine's are (relative L2, within a factor 2)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.flop_srp_spot())
    _say(f"case: flop SRP spot ({spot.num_nodes} nodes, hands {spot.num_hands}), f32, 20 iterations")
    gpu, cpu, ref = GpuSolver(spot), oracle.Oracle(spot), oracle.Oracle(spot, precision=64)
    pot = spot.starting_pot
    gpu.iterate(2)
    cpu.iterate(2)
    rg, sg = gpu.download_arenas()
    rc, sc = cpu.arenas()
    _say("  arenas after 2 iterations, entry by entry:")
    _close_arenas_chunked(rg, rc, pot)
    _close_arenas_chunked(sg, sc, 1.0)
    del rg, sg, rc, sc
    for it in (10, 20):
        gpu.iterate(it - gpu.iteration)
        cpu.iterate(it - cpu.iteration)
        eg, ec = gpu.exploitability(), cpu.exploitability()
        _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f} chips ({abs(eg - ec) / pot:.2e} of the pot apart)")
        assert abs(eg - ec) / pot < POT_TOL
    ref.iterate(20)
    rg, sg = gpu.download_arenas()
    gpu.close()
    rc, sc = cpu.arenas()
    del cpu
    r64, s64 = ref.arenas_f64()
    del ref
    for name, a, b, c in (("regrets", rg, rc, r64), ("strategy sums", sg, sc, s64)):
        for p in range(2):
            d_gc, d_g64, d_c64 = _rel_l2(a[p], b[p]), _rel_l2(a[p], c[p]), _rel_l2(b[p], c[p])
            _say(f"  {name} of player {p} after 20 iterations, relative L2: gpu-cpu {d_gc:.2e}, "
                 f"gpu-fp64 {d_g64:.2e}, cpu-fp64 {d_c64:.2e}")
            assert d_g64 <= 2.0 * d_c64 + 1e-6


@pytest.mark.timeout(1200)
def test_flop_srp_precision_forms_judged_by_fp64(built):
    """f32, f16 arenas, f16 arenas with f16 rows, f16 + BF16 and f16 + FP8
    showdown products, each run
    to 0.3% pot (exploitability checked every 10 iterations on the GPU), and
    the reference's i16/u16 arenas at a fixed stop (they floor above the
    target on this spot): the FP64 oracle's best response judges each final
    average strategy, and must agree with the GPU's figure within 0.05% of
    the pot (best_response.rs:255-287 in f64)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.flop_srp_spot())
    pot = spot.starting_pot
    o64 = oracle.Oracle(spot, precision=64)
    for storage, showdown, cap in (("f32", "exact", 600), ("half", "exact", 600), ("half_rows", "exact", 600),
                                   ("half", "bf16", 600), ("half", "fp8", 700), ("compressed", "exact", 400)):
        _say(f"case: flop SRP, {storage} arenas, {showdown} showdown, to 0.3% pot (at most {cap} iterations)")
        g = GpuSolver(spot, storage=storage, showdown=showdown)
        e = None
        while g.iteration < cap:
            g.iterate(10)
            e = g.exploitability() / pot * 100.0
            if e <= 0.3:
                break
        it = g.iteration
        r, s = g.download_arenas()
        g.close()
        o64.set_arenas(r, s)
        del r, s
        e64 = o64.exploitability() / pot * 100.0
        _say(f"  {it} iterations: gpu {e:.4f}% pot, fp64 judge {e64:.4f}% pot")
        assert abs(e64 - e) <= 0.05
        if storage != "compressed":
            assert e <= 0.3 and e64 <= 0.35


def _th9h2c(raked):
    from gtopen_amd import SpotConfig, StreetSizing, TreeConfig

    s = StreetSizing.of
    tree = TreeConfig(starting_pot=60.0, effective_stack=200.0, oop=[s("50", "60"), s("50", ""), s("50", "")],
                      ip=[s("50", "60"), s("50", ""), s("50", "")], rake_pct=0.05 if raked else 0.0,
                      rake_cap=3.0 if raked else 0.0)
    return SpotConfig("Th9h2c", "QQ,JJ,TT,99,AhKh,AsKs,87s", "AA,KK,AQs,JTs,T9s", tree)


def test_exploitability_matches_cpu_with_and_without_rake(built):
    """The reference's gpu_exploitability_matches_cpu (crates/solver/tests/
    gpu.rs:133-165): 100 GPU iterations on Th9h2c, then the GPU's
    exploitability and the CPU's on the downloaded arenas agree within 0.02%
    of the pot, unraked and with 5% rake capped at 3 (the raked formula
    ((br0 - v0) + (br1 - v1)) / 2, best_response.rs:271-284)."""
    import oracle
    from gtopen_amd import GpuSolver, Spot

    for raked in (False, True):
        spot = Spot(_th9h2c(raked))
        _say(f"case: Th9h2c, rake {'5% cap 3' if raked else 'none'}, 100 iterations")
        g = GpuSolver(spot, isomorphism=False)
        g.iterate(100)
        r, s = g.download_arenas()
        e_gpu = g.exploitability()
        cpu = oracle.Oracle(spot, isomorphism=False)
        cpu.set_arenas(r, s)
        e_cpu = cpu.exploitability()
        diff = abs(e_cpu - e_gpu) / 60.0 * 100.0
        _say(f"  gpu {e_gpu:.6f} cpu {e_cpu:.6f} chips ({diff:.2e}% pot apart)")
        assert diff < 0.02
        # the same on the GPU's own f64 root dot: the CPU best response and
        # average values at the root (the terms of the raked formula)
        for p in range(2):
            for mode in (0, 1):
                vg = g.root_cfv(p, mode).astype(np.float64)
                vc = cpu.root_cfv(p, mode)
                assert np.abs(vg - vc).max() <= 1e-5 * 60.0 + 1e-5 * np.abs(vc).max()
        g.close()


def test_raked_trajectories_match_oracle(built):
    """Raked DCFR (fold and showdown payoffs net of rake, tree.rs:580-601,
    761-777): GPU and oracle trajectories within 1e-4 of the pot, checked
    every 10 iterations for 50, on Th9h2c and on the turn spot with 5% rake
    capped at 3."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, SpotConfig, TreeConfig, workloads

    base = workloads.turn_spot()
    t = base.tree
    turn = SpotConfig(base.board, base.range_oop, base.range_ip,
                      TreeConfig(starting_pot=t.starting_pot, effective_stack=t.effective_stack, oop=t.oop, ip=t.ip,
                                 rake_pct=0.05, rake_cap=3.0))
    for name, cfg in (("Th9h2c", _th9h2c(True)), ("turn", turn)):
        spot = Spot(cfg)
        _say(f"case: {name} spot with 5% rake capped at 3, 50 iterations, check every 10")
        gpu, cpu = GpuSolver(spot), oracle.Oracle(spot)
        pot = spot.starting_pot
        for _ in range(5):
            gpu.iterate(10)
            cpu.iterate(10)
            eg, ec = gpu.exploitability(), cpu.exploitability()
            _say(f"  iter {gpu.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
            assert abs(eg - ec) <= POT_TOL * pot
        gpu.close()


@pytest.mark.timeout(900)
def test_batched_flops_through_the_batch_path_match_oracle(built):
    """Report flops of the batched-flops workload (1176-combo Beta(0.7,0.7)
    ranges, random bet menus, no raises) advanced together by the batch:
    20 iterations against the oracle, exploitability
    within 1e-4 of the pot at 10 and 20."""
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.batch import Batch

    cfgs = workloads.batched_flops(n=184)
    picks = [1, 3, 5]  # three tree shapes: {125}, {125, a}, {25}
    spots = [Spot(cfgs[i]) for i in picks]
    _say("case: batched flops " + ", ".join(f"{cfgs[i].board} ({s.num_nodes} nodes)" for i, s in zip(picks, spots)))
    gpus = [GpuSolver(s) for s in spots]
    b = Batch(gpus)
    cpus = [oracle.Oracle(s) for s in spots]
    for _ in range(2):
        b.iterate(10)
        for c in cpus:
            c.iterate(10)
        for i, (g, c, s) in enumerate(zip(gpus, cpus, spots)):
            eg, ec = g.exploitability(), c.exploitability()
            _say(f"  flop {cfgs[picks[i]].board} iter {g.iteration}: gpu {eg:.6f} cpu {ec:.6f}")
            assert g.iteration == c.iteration
            assert abs(eg - ec) <= POT_TOL * s.starting_pot
    b.close()
    for g in gpus:
        g.close()


def test_run_to_target(built):
    """gto_hip_run (Solver::run, cfr.rs:194-226): checks every check_every
    iterations, stops at the target or at max_iterations, the callback sees
    each check and may stop the run; the trajectory equals a manual loop's."""
    from gtopen_amd import GpuSolver, Spot, workloads

    spot = Spot(workloads.river_spot())
    pot = spot.starting_pot
    _say("case: river spot, run(max 400, target 1% pot, check every 20)")
    seen = []
    g = GpuSolver(spot)
    last = g.run(max_iterations=400, target_exploit_pct=1.0, check_every=20, on_progress=seen.append)
    assert seen and seen[-1] == last
    assert [p.iteration for p in seen] == list(range(20, 20 * len(seen) + 1, 20))
    assert last.exploit_pct_pot <= 1.0 and all(p.exploit_pct_pot > 1.0 for p in seen[:-1])
    assert last.iteration == g.iteration
    assert 0.0 < last.gpu_secs <= last.elapsed_secs
    ref = GpuSolver(spot)
    for p in seen:
        ref.iterate(p.iteration - ref.iteration)
        assert ref.exploitability() == p.exploit_chips
        assert p.exploit_pct_pot == pytest.approx(p.exploit_chips / pot * 100.0, rel=1e-12)
    # the callback stops the run; max_iterations ends it on a check
    g2 = GpuSolver(spot)
    last2 = g2.run(max_iterations=1000, target_exploit_pct=0.0, check_every=7, on_progress=lambda p: p.iteration < 21)
    assert last2.iteration == 21 == g2.iteration
    g3 = GpuSolver(spot)
    last3 = g3.run(max_iterations=25, target_exploit_pct=0.0, check_every=10)
    assert last3.iteration == 25 == g3.iteration
    for x in (g, ref, g2, g3):
        x.close()


def test_solve_cli_output_parses(built, tmp_path):
    """python -m gtopen_amd.cli prints solve-cli's progress and done lines
    (solve_cli.rs:432-450); the reference's chart script regex
    (bench/make_chart.py:34-36) reads every progress line."""
    from gtopen_amd import workloads

    path = tmp_path / "river.json"
    path.write_text(workloads.river_spot().to_json())
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    _say("case: solve-cli driver on the river spot, 200 iterations, target 0.5% pot")
    out = subprocess.run([sys.executable, "-m", "gtopen_amd.cli", str(path), "200", "0.5"], cwd=root,
                         capture_output=True, text=True, timeout=300, check=True).stdout
    lines = out.strip().splitlines()
    _say("  " + "\n  ".join(lines))
    assert lines[0].startswith("tree built in ") and "storage F32, algo Dcfr" in lines[0]
    it_re = r"iter\s+(\d+)\s+exploitability\s+([\d.]+) chips \(([\d.]+)% pot\)\s+\[([\d.]+)s(?:, gpu ([\d.]+)s)?\]"
    iters = [re.match(it_re, ln) for ln in lines[1:-1]]
    assert iters and all(iters)
    assert [int(m.group(1)) for m in iters] == list(range(10, 10 * len(iters) + 1, 10))
    assert all(m.group(5) is not None for m in iters)
    assert re.match(r"done: iter (\d+) exploitability ([\d.]+)% pot in ([\d.]+)s \(gpu kernels ([\d.]+)s = (\d+) "
                    r"ms/iter\)  peak_rss (\d+) MB", lines[-1])
