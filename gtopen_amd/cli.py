"""solve-cli on the MI355X engine: solve a spot from a SpotConfig JSON file
and print progress in the reference's format (crates/solver/src/bin/
solve_cli.rs:30-120 and run_gpu 382-455), so tools that parse solve-cli
output (bench/make_chart.py:34-36) read these runs unchanged.

    python -m gtopen_amd.cli <config.json> [max_iterations] [target_exploit_pct]

Environment (solve_cli.rs:3-11): SOLVER_STORAGE=f32|i16|half|half_rows
(arena storage on the device; i16 is the reference's compressed form,
half_rows f16 arenas with f16 rows in the iterations; default f32), SOLVER_ALGO=dcfr|cfr+|pcfr+, SOLVER_ISO=0 disables suit isomorphism.
Exploitability is checked every 10 iterations, as solve-cli's GPU path does.
"""
import os
import sys
import time

STORAGE_NAMES = {"f32": "F32", "compressed": "Compressed", "half": "Half", "half_rows": "HalfRows"}
ALGO_NAMES = {0: "Dcfr", 1: "CfrPlus", 2: "PcfrPlus"}


def peak_rss_mb():
    """VmHWM of this process in MB (solve_cli.rs:17-28)."""
    try:
        for line in open("/proc/self/status"):
            if line.startswith("VmHWM:"):
                return float(line.split()[1]) / 1024.0
    except OSError:
        pass
    return float("nan")


def header_line(secs, nodes, action_nodes, hands, arena_bytes, storage, algo):
    return (f"tree built in {secs:.2f}s: {nodes} nodes ({action_nodes} action nodes), hands {hands[0]}/{hands[1]}, "
            f"arenas {arena_bytes / 1e6:.1f} MB, storage {storage}, algo {algo}")


def progress_line(p):
    """The per-check line of run_gpu (solve_cli.rs:432-439)."""
    return (f"iter {p.iteration:5d}  exploitability {p.exploit_chips:8.4f} chips ({p.exploit_pct_pot:.3f}% pot)  "
            f"[{p.elapsed_secs:.1f}s, gpu {p.gpu_secs:.1f}s]")


def done_line(p, rss_mb):
    """The final line of run_gpu (solve_cli.rs:442-450)."""
    ms = p.gpu_secs * 1000.0 / p.iteration if p.iteration else float("nan")
    return (f"done: iter {p.iteration} exploitability {p.exploit_pct_pot:.3f}% pot in {p.elapsed_secs:.1f}s "
            f"(gpu kernels {p.gpu_secs:.2f}s = {ms:.0f} ms/iter)  peak_rss {rss_mb:.0f} MB")


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python -m gtopen_amd.cli <config.json> [max_iterations] [target_exploit_pct]", file=sys.stderr)
        return 1
    from . import _lib as L
    from .solver import GpuSolver
    from .spot import Spot

    storage = {"f32": "f32", "i16": "compressed", "compressed": "compressed", "half": "half", "f16": "half",
               "half_rows": "half_rows"}[
        os.environ.get("SOLVER_STORAGE", "f32")]
    algo = os.environ.get("SOLVER_ALGO", "dcfr")
    iso = os.environ.get("SOLVER_ISO", "1") != "0"
    max_iterations = int(argv[1]) if len(argv) > 1 else 300
    target = float(argv[2]) if len(argv) > 2 else 0.3

    t0 = time.perf_counter()
    spot = Spot(open(argv[0]).read())
    code = L.algo_code(algo)
    print(header_line(time.perf_counter() - t0, spot.num_nodes, spot.num_action_nodes, spot.num_hands,
                      spot.arena_bytes(L.STORAGE[storage]), STORAGE_NAMES[storage], ALGO_NAMES[code]), flush=True)
    gpu = GpuSolver(spot, storage=storage, algo=algo, isomorphism=iso)
    last = gpu.run(max_iterations=max_iterations, target_exploit_pct=target, check_every=10,
                   on_progress=lambda p: print(progress_line(p), flush=True))
    print(done_line(last, peak_rss_mb()), flush=True)
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
