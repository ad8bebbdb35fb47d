#!/usr/bin/env python3
"""DCFR benchmark on one MI355X: flop SRP spot (1,445,723-node tree).

A step is one DCFR iteration (both traversers, alternating updates) over the
whole tree, with the arenas resident in HBM. Prints one JSON line:
iterations/s (value), seconds to 0.3% pot exploitability, the roofline of the
dominant kernel, the CPU oracle's rate on a bounded sample, every precision
form and algorithm with its exploitability judged by the FP64 oracle, and the
other workloads (river, turn, a batched-flops sample, the differentiable game
value).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    """CPU threads this process may run on (the box's share, not the host's)."""
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_info():
    """Model name, physical cores and logical CPUs from lscpu."""
    info = {"cpu_model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    kv = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            kv[k.strip()] = v.strip()
    info["cpu_model"] = kv.get("Model name")
    try:
        info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (KeyError, ValueError):
        pass
    return info


DTYPE = {"f32": "f32", "compressed": "i16/u16 arenas, f32 math", "half": "f16 arenas, f32 math",
         "half_rows": "f16 arenas and rows, f32 math"}


def traffic_profile(args, storage=None):
    """The committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of this
    workload (profiles/<config>_<storage>_traffic.json), if any."""
    path = os.path.join(ROOT, "profiles", f"{args.config}_{storage or args.storage}_traffic.json")
    return json.load(open(path)) if os.path.exists(path) else None


def iteration_traffic_of(args, storage, D):
    """HBM bytes of a whole DCFR iteration from the committed counter
    profile of this workload and storage form, against the algorithmic arena
    bytes (20 B per entry with f32 arenas, 10 B with 16-bit arenas)."""
    tp = traffic_profile(args, storage)
    if tp is None:
        return None
    alg = (20.0 if storage == "f32" else 10.0) * D
    tb = sum(k["read_bytes_per_iter"] + k["write_bytes_per_iter"] for k in tp["kernels"].values())
    return {"bytes": round(tb), "algorithmic_bytes": alg, "ratio": round(tb / alg, 3),
            "source": f"profiles/{args.config}_{storage}_traffic.json (all kernels of the profiled run, per iteration)"}


def measured_traffic(args, kind):
    """HBM bytes per iteration of the kernel kind from the traffic summary."""
    prof = traffic_profile(args)
    if prof is None:
        return None
    store = {"f32": 0, "compressed": 1, "half": 2, "half_rows": 2}[args.storage]
    f16 = args.storage == "half_rows"  # the kernels' f16-row template argument (a flag, or a reach/value mask)
    if kind == "river":
        names = [k for k in prof["kernels"] if k.startswith("k_river<") and k.endswith(f", {store}>")]
    elif kind == "up_action":
        names = [f"k_up_action{v}<0, {store}, {x}>" for v in ("2", "") for x in (("3", "true") if f16 else ("0", "false"))]
    elif kind == "down_action":
        names = ["k_down_%s<%d, %s>" % (v, {0: 0, 1: 2, 2: 5}[store], "true" if f16 else "false")
                 for v in ("level", "action")]
    else:
        names = []
    key = next((k for k in names if k in prof["kernels"]), None)
    k = prof["kernels"].get(key)
    if not k:
        return None
    return round(k["read_bytes_per_iter"] + k["write_bytes_per_iter"])


class Fp64Judge:
    """The FP64 oracle's best-response exploitability (best_response.rs:
    255-287 in f64) of a GPU engine's final average strategy: the checker,
    run outside every timed region."""

    def __init__(self, spot):
        import oracle

        oracle.set_threads(host_threads())
        self.o = oracle.Oracle(spot, precision=64)
        self.pot = spot.starting_pot

    def __call__(self, g):
        r, s = g.download_arenas()
        self.o.set_arenas(r, s)
        del r, s
        return self.o.exploitability() / self.pot * 100.0


def time_to_target(spot, storage, target, cap, showdown="exact", sched=None, algo="dcfr", judge=None):
    """Wall time of a fresh engine to `target` % pot, best-response check
    every 10 iterations, at most `cap` iterations: (seconds, iterations,
    %pot, FP64-judged %pot or None). sched: the autotuner's choice on this
    spot, applied to the fresh engine; judge: run on the final engine after
    the clock stops."""
    from gtopen_amd import GpuSolver

    g = GpuSolver(spot, storage=storage, showdown=showdown, algo=algo)
    if sched is not None:
        g.set_schedule(sched["chosen"]["show_nodes"], sched["chosen"]["per_level"], sched["chosen"]["river"])
    t0 = time.perf_counter()
    e = None
    while g.iteration < cap:
        g.iterate(10)
        e = g.exploitability() / spot.starting_pot * 100.0
        if e <= target:
            break
    dt = time.perf_counter() - t0
    its = g.iteration
    e64 = judge(g) if judge is not None else None
    g.close()
    return dt, its, e, e64


def gpu_rate(spot, steps, warmup, storage="f32", showdown="exact", autotune=0):
    """(ms per iteration, iterations/s, arena entries D) of graph-replayed
    iterations; autotune > 0: the autotuner's schedule first (that many
    iterations per candidate, before the warm-up)."""
    from gtopen_amd import GpuSolver

    g = GpuSolver(spot, storage=storage, showdown=showdown)
    if autotune > 0:
        g.autotune(autotune)
    g.iterate(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    g.iterate(steps)
    g.synchronize()
    dt = time.perf_counter() - t0
    D = g.reached_entries
    g.close()
    return dt * 1e3 / steps, steps / dt, D


def roofline_of(prof, steps, D, Dr, eb, args=None):
    """Roofline of the dominant kernel kind of a profile: algorithmic arena
    bytes of that kind per iteration over its summed launch time. Dr: the
    arena entries inside river subtrees, which the fused river kernel
    streams when it runs (the level kernels then cover D - Dr)."""
    kinds = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max(kinds, key=lambda k: kinds[k][0])
    dom_ms = kinds[dom][0] / steps
    if "river" not in kinds:
        Dr = 0
    # regret/strategy arena entries: the update reads+writes R and S (16 B
    # per f32 entry per iteration, both sweeps), the down sweep reads R once
    # more (4 B): 20 B per entry per iteration in total, all of it in the
    # fused river kernel for river entries
    alg = {"up_action": 4.0 * eb * (D - Dr), "down_action": eb * (D - Dr), "river": 5.0 * eb * Dr}.get(dom)
    split = {k: round(v[0] / steps, 4) for k, v in kinds.items()}
    if alg is None:
        return None, split
    achieved = alg / (dom_ms * 1e-3) / 1e9
    r = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(args, dom) if args else None,
         "basis": "bytes and time summed over one iteration's launches of the kernel",
         "avg_launch_us": round(kinds[dom][0] * 1e3 / kinds[dom][1], 2), "algorithmic_bytes_per_iter": alg}
    return r, split


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="flop_srp", choices=["river", "turn", "flop_srp", "batched_flops"])
    ap.add_argument("--flops", type=int, default=184, help="batched_flops: number of seeded flops")
    ap.add_argument("--flop-iters", type=int, default=300, help="batched_flops: DCFR iterations per flop")
    ap.add_argument("--target", type=float, default=0.3, help="exploitability target, %% of pot")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-target", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the precision forms and algorithms")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 oracle's judgement of each result")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the river, turn, batched-flops sample and game-value lines")
    ap.add_argument("--sample-flops", type=int, default=16, help="flops in the batched-flops sample line")
    ap.add_argument("--autotune", type=int, default=2,
                    help="iterations per candidate launch schedule for the autotuner (0: default schedule)")
    ap.add_argument("--showdown", default="exact", choices=["exact", "sequential", "bf16", "fp8"],
                    help="showdown evaluation: exact f64 prefix sums, the same with the fold and showdown sums in "
                         "the CPU engine's order (sequential), or the BF16 / FP8 matrix-core products")
    ap.add_argument("--storage", default="f32", choices=["f32", "compressed", "half", "half_rows"])
    ap.add_argument("--river", default="auto", choices=["auto", "fused", "level"],
                    help="with --autotune 0: the fused river sweep or the level sweeps (auto: the engine default)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB mebibytes (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch

        n = args.copy_probe << 18  # float32 elements
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"bench: copy probe {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported "
              "(the tree would shard by turn card across GPUs; not built)", file=sys.stderr)
        sys.exit(2)

    if args.config == "batched_flops":
        return bench_batched(args)

    from gtopen_amd import GpuSolver, Spot, workloads

    cfg = workloads.CONFIGS[args.config]()
    t0 = time.time()
    spot = Spot(cfg)
    log(f"bench: {args.config} spot built in {time.time() - t0:.2f}s: {spot.num_nodes} nodes, "
        f"{spot.num_action_nodes} action nodes, hands {spot.num_hands}")
    pot = spot.starting_pot

    gpu = GpuSolver(spot, storage=args.storage, showdown=args.showdown)
    D, Dr = gpu.reached_entries, gpu.river_entries
    if args.river != "auto":
        gpu.set_schedule(16, False, args.river == "fused")
    gpu.iterate(args.warmup)
    gpu.synchronize()
    log(f"bench: warmup {args.warmup} done")
    sched = None
    if args.autotune > 0:
        chosen, cand = gpu.autotune(args.autotune)
        sched = {"chosen": chosen, "candidates_ms_per_iter": [round(x, 3) for x in cand],
                 "candidates": "level sweeps: show_nodes 16/8/4 x terminals all-levels-first / per-level; "
                               "fused river sweep"}
        log(f"bench: autotuner chose {chosen}; candidates ms/iter {sched['candidates_ms_per_iter']}")
    t0 = time.perf_counter()
    gpu.iterate(args.steps)
    gpu.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = dt * 1e3 / args.steps
    its = args.steps / dt
    e_after = gpu.exploitability()
    log(f"bench: {args.steps} steps in {dt:.3f}s = {ms_per_step:.3f} ms/iter, "
        f"exploitability after {gpu.iteration} iterations {e_after / pot * 100:.4f}% pot")

    # per-kernel split with HIP events on the engine's stream, over the same
    # number of steps (eager launches, one event pair per launch)
    prof = gpu.profile(args.steps)
    eb = 4.0 if args.storage == "f32" else 2.0  # bytes per arena entry
    roofline, split = roofline_of(prof, args.steps, D, Dr, eb, args)
    log("bench: kernel split (ms/iter): " + ", ".join(f"{k} {v:.3f}" for k, v in split.items()))
    whole = 5.0 * eb * D / (ms_per_step * 1e-3) / 1e9
    gpu.close()
    # HBM bytes of a whole iteration from the committed counter profile
    iteration_traffic = iteration_traffic_of(args, args.storage, D)

    judge = None
    if not args.no_target and not args.no_fp64:
        t1 = time.perf_counter()
        judge = Fp64Judge(spot)
        log(f"bench: FP64 oracle judge ready in {time.perf_counter() - t1:.1f}s")

    sec_to_target, iters_to_target, e_final, e64_final = None, None, None, None
    if not args.no_target:
        sec_to_target, iters_to_target, e_final, e64_final = time_to_target(
            spot, args.storage, args.target, 2000, args.showdown, sched, judge=judge)
        log(f"bench: {iters_to_target} iterations to {e_final:.4f}% pot in {sec_to_target:.3f}s (checks every 10)"
            + ("" if e64_final is None else f"; FP64 judge {e64_final:.4f}% pot"))

    # the precision forms on the same spot: f16 arenas (this build's format)
    # with f32 or f16 rows and the exact showdown, with the BF16 and FP8
    # showdowns, and the reference's i16/u16
    # arenas; each final strategy judged by the FP64 oracle
    mixed = None
    if not args.no_target and not args.no_mixed and args.storage == "f32":
        mixed = {}
        for key, st, sd, cap in (("half", "half", "exact", 2000), ("half_rows", "half_rows", "exact", 2000),
                                 ("half_bf16_showdown", "half", "bf16", 2000),
                                 ("half_fp8_showdown", "half", "fp8", 1000), ("compressed", "compressed", "exact", 800)):
            ms, _, _ = gpu_rate(spot, args.steps, args.warmup, st, sd)
            sec, its_t, e_t, e64 = time_to_target(spot, st, args.target, cap, sd, sched, judge=judge)
            reached = e_t <= args.target
            mixed[key] = {"arenas": {"half": "f16 x per-node power of two", "half_rows": "f16 x per-node power of two",
                                     "compressed": "i16/u16 (reference)"}[st],
                          "rows": "f16 x per-row power of two" if st == "half_rows" else "f32",
                          "showdown": {"exact": "f64 prefix sums", "bf16": "BF16 matrix-core products, f32 accumulate",
                                       "fp8": "FP8 (e4m3, per-node 2^k scale) matrix-core products, f32 accumulate"}[sd],
                          "ms_per_step": round(ms, 4), "iters_per_s": round(1e3 / ms, 3),
                          "sec_to_target": round(sec, 3) if reached else None,
                          "iters_to_target": its_t if reached else None,
                          "exploitability_pct_at_stop": round(e_t, 4),
                          "iterations_at_stop": its_t,
                          "fp64_exploitability_pct": None if e64 is None else round(e64, 4),
                          "iteration_traffic": iteration_traffic_of(args, st, D) if sd == "exact" else None}
            log(f"bench: {key}: {ms:.3f} ms/iter; " + (
                f"{its_t} iterations to {e_t:.4f}% pot in {sec:.3f}s" if reached
                else f"target not reached: {e_t:.4f}% pot after {its_t} iterations")
                + ("" if e64 is None else f"; FP64 judge {e64:.4f}% pot"))

    # f32 with the fold and showdown f64 sums in the CPU engine's order
    # (bit-identical to the CPU engine; the tests check every arena entry)
    seq = None
    if not args.no_target and not args.no_mixed and args.storage == "f32":
        ms, _, _ = gpu_rate(spot, args.steps, args.warmup, "f32", "sequential")
        sec, its_t, e_t, e64 = time_to_target(spot, "f32", args.target, 2000, "sequential", sched, judge=judge)
        reached = e_t <= args.target
        seq = {"showdown": "f64 sums in the CPU engine's order where the association could change a bit",
               "ms_per_step": round(ms, 4), "iters_per_s": round(1e3 / ms, 3),
               "sec_to_target": round(sec, 3) if reached else None,
               "iters_to_target": its_t if reached else None,
               "exploitability_pct_at_stop": round(e_t, 4),
               "fp64_exploitability_pct": None if e64 is None else round(e64, 4)}
        log(f"bench: sequential sums: {ms:.3f} ms/iter; " + (
            f"{its_t} iterations to {e_t:.4f}% pot in {sec:.3f}s" if reached
            else f"target not reached: {e_t:.4f}% pot after {its_t} iterations")
            + ("" if e64 is None else f"; FP64 judge {e64:.4f}% pot"))

    # the other algorithms of the solver crate on the same spot (f32 arenas)
    algos = None
    if not args.no_target and not args.no_mixed and args.storage == "f32":
        algos = {}
        for name in ("cfr+", "pcfr+"):
            sec, its_t, e_t, e64 = time_to_target(spot, "f32", args.target, 1000, "exact", sched, name, judge=judge)
            reached = e_t <= args.target
            algos[name] = {"sec_to_target": round(sec, 3) if reached else None,
                           "iters_to_target": its_t if reached else None,
                           "exploitability_pct_at_stop": round(e_t, 4),
                           "fp64_exploitability_pct": None if e64 is None else round(e64, 4)}
            log(f"bench: {name}: " + (f"{its_t} iterations to {e_t:.4f}% pot in {sec:.3f}s" if reached
                                      else f"target not reached: {e_t:.4f}% pot after {its_t} iterations")
                + ("" if e64 is None else f"; FP64 judge {e64:.4f}% pot"))
    judge = None

    cpu = None
    if not args.no_cpu:
        import oracle

        threads = host_threads()
        oracle.set_threads(threads)
        o = oracle.Oracle(spot)
        t0 = time.perf_counter()
        n = 0
        while True:
            o.iterate(1)
            n += 1
            if time.perf_counter() - t0 >= args.cpu_seconds or n >= 50:
                break
        cdt = time.perf_counter() - t0
        del o
        info = cpu_info()
        cpu = {"value": round(n / cdt, 4), "unit": "DCFR iters/s", "cores": threads, "kind": "port",
               "sample": f"first {n} DCFR iterations of the same {args.config} spot, f32 arenas, "
                         f"C++ oracle (restated reference CPU engine) on {threads} threads",
               "cpu_model": info["cpu_model"], "physical_cores": info["physical_cores"],
               "logical_cpus": info["logical_cpus"]}
        if seq is not None and seq["iters_to_target"]:
            # the sequential-sums run is bit-identical to the CPU engine, so
            # its count is the CPU's own; only the time per iteration is
            # extrapolated
            its_cpu = seq["iters_to_target"]
            cpu.update({"sec_to_target": round(its_cpu * cdt / n, 1), "iters_to_target": its_cpu,
                        "extrapolated": True,
                        "method": f"the CPU engine's iteration count to {args.target}% pot ({its_cpu}, checks every "
                                  "10: the GPU's sequential-sums run, bit-identical to the CPU engine) times the "
                                  f"oracle's mean seconds per iteration over its first {n}; later iterations prune "
                                  "more and run faster, so this overstates the CPU time"})
        elif iters_to_target:
            cpu.update({"sec_to_target": round(iters_to_target * cdt / n, 1), "extrapolated": True,
                        "method": f"the GPU's f32 iteration count to {args.target}% pot ({iters_to_target}, the "
                                  "same arithmetic as the oracle's up to the order of some f64 sums) times the "
                                  f"oracle's mean seconds per iteration over its first {n}; later iterations "
                                  "prune more and run faster, so this overstates the CPU time"})
        log(f"bench: cpu oracle {n} iterations in {cdt:.2f}s on {threads} threads ({info['cpu_model']}, "
            f"{info['physical_cores']} physical cores)")

    configs = None
    if not args.no_configs and args.config == "flop_srp":
        configs = other_configs(args)

    out = {
        "metric": f"DCFR iters/s on the {args.config} spot ({args.storage} arenas); sec to "
                  f"{args.target}%-pot exploitability and the mixed-precision forms alongside",
        "value": round(its, 3),
        "unit": "DCFR iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.storage]
                 + {"exact": "", "sequential": ", f64 terminal sums in the CPU engine's order"}.get(
                     args.showdown, f", {args.showdown} showdown products"),
        "data": "synthetic",
        "config": {"workload": f"{args.config} spot, seeded (gtopen_amd.workloads)", "nodes": spot.num_nodes,
                   "hands": list(spot.num_hands), "arena_entries_D": D, "river_arena_entries": Dr,
                   "parallelism": "single GPU"},
        "sec_to_target": None if sec_to_target is None else round(sec_to_target, 3),
        "iters_to_target": iters_to_target,
        "target_pct_pot": args.target,
        "exploitability_pct_at_target": None if e_final is None else round(e_final, 4),
        "fp64_exploitability_pct": None if e64_final is None else round(e64_final, 4),
        "whole_iteration_arena_GBps": round(whole, 1),
        "iteration_traffic": iteration_traffic,
        "kernel_ms_per_iter": split,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "sequential_sums": seq,
        "algorithms": algos,
        "schedule": sched,
        "configs": configs,
    }
    print(json.dumps(out), flush=True)


def other_configs(args):
    """The other workloads, one entry each: the river and turn spots (GPU
    iterations/s, arena bytes/s), a batched-flops sample through the report
    runner (flops/hour with its arena roofline), and the differentiable game
    value on the turn spot (value + gradient evaluations/s)."""
    from gtopen_amd import Spot, workloads

    out = {}
    for name, steps in (("river", 500), ("turn", 200)):
        spot = Spot(workloads.CONFIGS[name]())
        ms, ips, D = gpu_rate(spot, steps, 10, autotune=args.autotune)
        gbs = 20.0 * D / (ms * 1e-3) / 1e9
        out[name] = {"iters_per_s": round(ips, 2), "ms_per_step": round(ms, 4), "arena_entries_D": D,
                     "nodes": spot.num_nodes, "hands": list(spot.num_hands),
                     "arena_GBps": round(gbs, 1), "arena_frac_of_peak": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"bench: {name} spot: {ms:.4f} ms/iter ({ips:.1f} iters/s), {gbs:.1f} GB/s of arena bytes")
    out["batched_flops_sample"] = batched_sample(args)
    out["game_value"] = game_value_rate()
    return out


def batched_sample(args):
    """The first `sample_flops` flops of the batched-flops workload, 300
    iterations each through the report runner (waves of engines advanced by
    one parent-graph launch per iteration); arena roofline = sum over flops
    of 20 B (f32) per reached arena entry per iteration, over the wall time."""
    from gtopen_amd import workloads
    from gtopen_amd.reports import run_report

    cfgs = workloads.batched_flops(n=184)[:args.sample_flops]
    t0 = time.perf_counter()
    rep = run_report([(c, 1) for c in cfgs], target=0.0, max_iterations=args.flop_iters,
                     check_every=args.flop_iters, name="bench-sample", autotune=0, storage="f32")
    dt = time.perf_counter() - t0
    alg = sum(20.0 * r["arena_entries"] * r["iterations"] for r in rep["flops"])
    gbs = alg / dt / 1e9
    res = {"flops": len(rep["flops"]), "iterations_per_flop": args.flop_iters, "seconds": round(dt, 2),
           "flops_per_hour": round(len(rep["flops"]) / dt * 3600.0, 1), "waves": rep["waves"],
           "arena_bytes": alg,
           "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(gbs / HBM_PEAK_GBS, 4),
                        "basis": "20 B per reached f32 arena entry per iteration, summed over flops, over the "
                                 "report's wall time (spot builds, engine creation and checks included)"},
           "mean_exploit_pct": round(sum(r["exploit_pct"] for r in rep["flops"]) / len(rep["flops"]), 4)}
    log(f"bench: batched-flops sample: {res['flops']} flops x {args.flop_iters} iterations in {dt:.1f}s "
        f"({res['flops_per_hour']} flops/hour, {gbs:.1f} GB/s of arena bytes)")
    return res


def game_value_rate():
    """Config 5: the turn spot solved for 500 iterations, strategies frozen;
    64 seeded range perturbations N(0, 0.02^2) clipped to [0, 1]; value and
    gradient per perturbation (both ranges; and with every node's strategy
    gradient as well)."""
    import numpy as np

    from gtopen_amd import GameValue, GpuSolver, Spot, workloads
    from gtopen_amd.game_value import behavioural_strategy

    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot)
    g.iterate(500)
    _, strat = g.download_arenas()
    g.close()
    gv = GameValue(spot)
    gv.set_strategy(*behavioural_strategy(spot, strat))
    rng = np.random.default_rng(2024)
    w = [spot.weights[q].astype(np.float64) for q in range(2)]
    ws = [[np.clip(w[q] + rng.normal(0.0, 0.02, size=w[q].shape), 0.0, 1.0).astype(np.float32) for q in range(2)]
          for _ in range(64)]
    res = {}
    for key, gs, dev in (("ranges", False, False), ("ranges_and_strategies", True, False),
                         ("ranges_and_strategies_on_gpu", True, True)):
        run = gv.value_and_grad_device if dev else gv.value_and_grad
        run(0, grad_strategy=gs)
        t0 = time.perf_counter()
        for wk in ws:
            gv.set_weights(*wk)
            run(0, grad_strategy=gs)
        dt = time.perf_counter() - t0
        res[key] = {"evals_per_s": round(64 / dt, 2), "ms_per_eval": round(dt * 1e3 / 64, 3)}
    n_sigma = spot.data_size[0] + spot.data_size[1]
    # per evaluation the strategy arenas are read (down pass) and the strategy
    # gradient written: 8 bytes per entry of both players' arenas
    gbs = 8.0 * n_sigma * res["ranges_and_strategies_on_gpu"]["evals_per_s"] / 1e9
    res["ranges_and_strategies_on_gpu"]["arena_GB_per_s"] = round(gbs, 1)
    res["workload"] = (f"turn spot ({spot.num_nodes} nodes), 500 DCFR iterations, 64 range perturbations; "
                       f"strategy gradient {n_sigma} floats, copied to host arrays, or left on the GPU as torch "
                       "tensors (_on_gpu; arena_GB_per_s: 8 bytes per strategy entry per evaluation, sigma read "
                       "and gradient written)")
    log(f"bench: game value + gradient: {res['ranges']['evals_per_s']} evals/s (ranges), "
        f"{res['ranges_and_strategies']['evals_per_s']} evals/s (ranges and strategies), "
        f"{res['ranges_and_strategies_on_gpu']['evals_per_s']} evals/s with the gradients left on the GPU "
        f"({gbs:.1f} GB/s of strategy arena bytes)")
    return res


def bench_batched(args):
    """Batched flops: seeded canonical flops with per-flop bet-size menus
    (uneven trees), full-range Beta(0.7,0.7) ranges, solved in waves of
    engines that iterate together (one parent-graph launch per iteration),
    a fixed number of DCFR iterations each; value = flops/hour."""
    from gtopen_amd import Spot, workloads
    from gtopen_amd.reports import run_report

    cfgs = workloads.batched_flops(n=args.flops)
    done = []

    def progress(row):
        done.append(row)
        log(f"bench: flop {len(done)}/{len(cfgs)} {row['board']} ({row['nodes']} nodes): "
            f"{row['exploit_pct']:.3f}% pot after {row['iterations']} iterations, {row['seconds']:.1f}s in its wave")

    t0 = time.perf_counter()
    rep = run_report([(c, 1) for c in cfgs], target=0.0, max_iterations=args.flop_iters,
                     check_every=args.flop_iters, progress=progress, name="bench", autotune=args.autotune,
                     storage=args.storage)
    for sc in rep.get("schedules", []):
        log(f"bench: wave {sc['wave']}: autotuner chose show_nodes {sc['show_nodes']}, per_level {sc['per_level']}; "
            f"candidates ms/iter {['%.1f' % x for x in sc['ms_per_iteration']]}")
    dt = time.perf_counter() - t0
    per_hour = len(rep["flops"]) / dt * 3600.0
    eb = 20.0 if args.storage == "f32" else 10.0
    alg = sum(eb * r["arena_entries"] * r["iterations"] for r in rep["flops"])
    gbs = alg / dt / 1e9
    log(f"bench: {len(rep['flops'])} flops x {args.flop_iters} iterations in {dt:.1f}s ({rep['waves']} waves), "
        f"{gbs:.1f} GB/s of arena bytes")
    cpu = None
    if not args.no_cpu:
        import oracle

        threads = host_threads()
        oracle.set_threads(threads)
        o = oracle.Oracle(Spot(cfgs[0]))
        t1 = time.perf_counter()
        n = 0
        while time.perf_counter() - t1 < args.cpu_seconds and n < args.flop_iters:
            o.iterate(1)
            n += 1
        cdt = time.perf_counter() - t1
        info = cpu_info()
        cpu = {"value": round(n / cdt / args.flop_iters * 3600.0, 2), "unit": "flops/hour", "cores": threads,
               "kind": "port",
               "sample": f"first {n} DCFR iterations of flop 1 ({cfgs[0].board}) on the C++ oracle, "
                         f"scaled to {args.flop_iters} iterations per flop",
               "cpu_model": info["cpu_model"], "physical_cores": info["physical_cores"],
               "logical_cpus": info["logical_cpus"]}
        log(f"bench: cpu oracle {n} iterations of flop 1 in {cdt:.2f}s on {threads} threads")
    out = {
        "metric": f"flops/hour, {len(cfgs)} batched flops x {args.flop_iters} DCFR iterations ({args.storage} arenas)",
        "value": round(per_hour, 2), "unit": "flops/hour", "n_gpus": 1, "steps": args.flop_iters, "warmup": 0,
        "ms_per_step": round(dt * 1e3 / max(1, args.flop_iters), 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.storage],
        "data": "synthetic",
        "config": {"workload": f"batched_flops: {len(cfgs)} seeded canonical flops, per-flop menus from "
                               "{25,33,50,75,125,a}, 1176-combo Beta(0.7,0.7) ranges",
                   "nodes_total": int(sum(r["nodes"] for r in rep["flops"])), "waves": rep["waves"],
                   "parallelism": "single GPU, the engines of a wave iterating together (each engine's iteration graph on its own stream)",
                   "schedules": [{k: sc[k] for k in ("wave", "show_nodes", "per_level")} for sc in rep.get("schedules", [])]},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "basis": f"{eb:.0f} B per reached arena entry per iteration, summed over flops, over the wall "
                              "time of the whole run"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
