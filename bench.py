#!/usr/bin/env python3
"""DCFR benchmark on one MI355X: flop SRP spot (1,445,723-node tree).

A step is one DCFR iteration (both traversers, alternating updates) over the
whole tree, with the arenas resident in HBM. Prints one JSON line:
iterations/s (value), seconds to 0.3% pot exploitability, the roofline of the
dominant kernel and the CPU oracle's rate on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(args, kind):
    """HBM bytes per iteration of the kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (profiles/), if any."""
    path = os.path.join(ROOT, "profiles", f"{args.config}_{args.storage}_traffic.json")
    if not os.path.exists(path):
        return None
    prof = json.load(open(path))
    store = {"f32": 0, "compressed": 1, "half": 2}[args.storage]
    key = {"up_action": "k_up_action<0, %d>" % store,
           "down_action": "k_down_action<%d>" % {0: 0, 1: 2, 2: 5}[store]}.get(kind)
    k = prof["kernels"].get(key)
    if not k:
        return None
    return round(k["read_bytes_per_iter"] + k["write_bytes_per_iter"])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="flop_srp", choices=["river", "turn", "flop_srp", "batched_flops"])
    ap.add_argument("--flops", type=int, default=184, help="batched_flops: number of seeded flops")
    ap.add_argument("--flop-iters", type=int, default=300, help="batched_flops: DCFR iterations per flop")
    ap.add_argument("--target", type=float, default=0.3, help="exploitability target, % of pot")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="CPU baseline sample budget")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-target", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the f16 and i16/u16 arena comparison")
    ap.add_argument("--autotune", type=int, default=2,
                    help="iterations per candidate launch schedule for the autotuner (0: default schedule)")
    ap.add_argument("--showdown", default="exact", choices=["exact", "bf16", "fp8"],
                    help="showdown evaluation: exact f64 prefix sums or the BF16 / FP8 matrix-core products")
    ap.add_argument("--storage", default="f32", choices=["f32", "compressed", "half"])
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
    D = gpu.reached_entries
    gpu.iterate(args.warmup)
    gpu.synchronize()
    log(f"bench: warmup {args.warmup} done")
    sched = None
    if args.autotune > 0:
        chosen, cand = gpu.autotune(args.autotune)
        sched = {"chosen": chosen, "candidates_ms_per_iter": [round(x, 3) for x in cand],
                 "candidates": "show_nodes 16/8/4 x terminals all-levels-first / per-level"}
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
    kinds = {k: v for k, v in prof.items() if v[1] > 0}
    dom = max(kinds, key=lambda k: kinds[k][0])
    dom_ms_iter = kinds[dom][0] / args.steps
    dom_launches = kinds[dom][1]
    log("bench: kernel split (ms/iter, launches/iter): " +
        ", ".join(f"{k} {v[0] / args.steps:.3f} ({v[1] // args.steps})" for k, v in kinds.items()))
    # algorithmic bytes: regret/strategy arena entries D; the update kernel
    # reads+writes R and S (16 B per entry per iteration, both sweeps), the
    # down sweep reads R once more (4 B): 20*D per iteration in total.
    eb = 4.0 if args.storage == "f32" else 2.0  # bytes per arena entry
    alg = {"up_action": 4.0 * eb * D, "down_action": eb * D}
    dom_bytes = alg.get(dom)
    roofline = None
    if dom_bytes is not None:
        achieved = dom_bytes / (dom_ms_iter * 1e-3) / 1e9
        roofline = {"bound": "hbm", "kernel": f"k_{dom}", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(args, dom),
                    "basis": "bytes and time summed over one iteration's launches of the kernel",
                    "avg_launch_us": round(kinds[dom][0] * 1e3 / dom_launches, 2),
                    "algorithmic_bytes_per_iter": dom_bytes}
    whole = 5.0 * eb * D / (ms_per_step * 1e-3) / 1e9

    sec_to_target, iters_to_target, e_final = None, None, None
    if not args.no_target:
        sec_to_target, iters_to_target, e_final = time_to_target(spot, args.storage, args.target, 2000,
                                                                 args.showdown, sched)
        log(f"bench: {iters_to_target} iterations to {e_final:.4f}% pot in {sec_to_target:.3f}s (checks every 10)")
    gpu.close()

    # the mixed-precision forms on the same spot: f16 arenas (this build's
    # format) and the reference's i16/u16 arenas
    mixed = None
    if not args.no_target and not args.no_mixed and args.storage == "f32":
        mixed = {}
        for key, st, sd, cap in (("half", "half", "exact", 2000), ("half_bf16_showdown", "half", "bf16", 2000),
                                 ("half_fp8_showdown", "half", "fp8", 1000), ("compressed", "compressed", "exact", 800)):
            g3 = GpuSolver(spot, storage=st, showdown=sd)
            g3.iterate(args.warmup)
            g3.synchronize()
            t0 = time.perf_counter()
            g3.iterate(args.steps)
            g3.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
            g3.close()
            sec, its_t, e_t = time_to_target(spot, st, args.target, cap, sd, sched)
            reached = e_t <= args.target
            mixed[key] = {"arenas": {"half": "f16 x per-node power of two", "compressed": "i16/u16 (reference)"}[st],
                          "showdown": {"exact": "f64 prefix sums", "bf16": "BF16 matrix-core products, f32 accumulate",
                                       "fp8": "FP8 (e4m3, per-node 2^k scale) matrix-core products, f32 accumulate"}[sd],
                          "ms_per_step": round(ms, 4), "iters_per_s": round(1e3 / ms, 3),
                         "sec_to_target": round(sec, 3) if reached else None,
                         "iters_to_target": its_t if reached else None,
                         "exploitability_pct_at_stop": round(e_t, 4)}
            log(f"bench: {key}: {ms:.3f} ms/iter; " + (
                f"{its_t} iterations to {e_t:.4f}% pot in {sec:.3f}s" if reached
                else f"target not reached: {e_t:.4f}% pot after {its_t} iterations"))

    # the other algorithms of the solver crate on the same spot (f32 arenas)
    algos = None
    if not args.no_target and not args.no_mixed and args.storage == "f32":
        algos = {}
        for name in ("cfr+", "pcfr+"):
            sec, its_t, e_t = time_to_target(spot, "f32", args.target, 1000, "exact", sched, name)
            reached = e_t <= args.target
            algos[name] = {"sec_to_target": round(sec, 3) if reached else None,
                           "iters_to_target": its_t if reached else None,
                           "exploitability_pct_at_stop": round(e_t, 4)}
            log(f"bench: {name}: " + (f"{its_t} iterations to {e_t:.4f}% pot in {sec:.3f}s" if reached
                                      else f"target not reached: {e_t:.4f}% pot after {its_t} iterations"))

    cpu = None
    if not args.no_cpu:
        import oracle

        threads = os.cpu_count() or 1
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
        cpu = {"value": round(n / cdt, 4), "unit": "DCFR iters/s", "cores": threads, "kind": "port",
               "sample": f"first {n} DCFR iterations of the same {args.config} spot, f32 arenas, "
                         f"C++ oracle (restated reference CPU engine) on {threads} threads"}
        log(f"bench: cpu oracle {n} iterations in {cdt:.2f}s on {threads} threads")

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
        "dtype": {"f32": "f32", "compressed": "i16/u16 arenas, f32 math", "half": "f16 arenas, f32 math"}[args.storage]
                 + ("" if args.showdown == "exact" else f", {args.showdown} showdown products"),
        "data": "synthetic",
        "config": {"workload": f"{args.config} spot, seeded (gtopen_amd.workloads)", "nodes": spot.num_nodes,
                   "hands": list(spot.num_hands), "arena_entries_D": D, "parallelism": "single GPU"},
        "sec_to_target": None if sec_to_target is None else round(sec_to_target, 3),
        "iters_to_target": iters_to_target,
        "target_pct_pot": args.target,
        "whole_iteration_arena_GBps": round(whole, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "mixed_precision": mixed,
        "algorithms": algos,
        "schedule": sched,
    }
    print(json.dumps(out), flush=True)


def time_to_target(spot, storage, target, cap, showdown="exact", sched=None, algo="dcfr"):
    """Wall time of a fresh engine to `target` % pot, best-response check
    every 10 iterations, at most `cap` iterations: (seconds, iterations, %pot).
    sched: the autotuner's choice on this spot, applied to the fresh engine."""
    from gtopen_amd import GpuSolver

    g = GpuSolver(spot, storage=storage, showdown=showdown, algo=algo)
    if sched is not None:
        g.set_schedule(sched["chosen"]["show_nodes"], sched["chosen"]["per_level"])
    t0 = time.perf_counter()
    e = None
    while g.iteration < cap:
        g.iterate(10)
        e = g.exploitability() / spot.starting_pot * 100.0
        if e <= target:
            break
    dt = time.perf_counter() - t0
    its = g.iteration
    g.close()
    return dt, its, e


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
    log(f"bench: {len(rep['flops'])} flops x {args.flop_iters} iterations in {dt:.1f}s ({rep['waves']} waves)")
    cpu = None
    if not args.no_cpu:
        import oracle

        threads = os.cpu_count() or 1
        oracle.set_threads(threads)
        o = oracle.Oracle(Spot(cfgs[0]))
        t1 = time.perf_counter()
        n = 0
        while time.perf_counter() - t1 < args.cpu_seconds and n < args.flop_iters:
            o.iterate(1)
            n += 1
        cdt = time.perf_counter() - t1
        cpu = {"value": round(n / cdt / args.flop_iters * 3600.0, 2), "unit": "flops/hour", "cores": threads,
               "kind": "port",
               "sample": f"first {n} DCFR iterations of flop 1 ({cfgs[0].board}) on the C++ oracle, "
                         f"scaled to {args.flop_iters} iterations per flop"}
        log(f"bench: cpu oracle {n} iterations of flop 1 in {cdt:.2f}s on {threads} threads")
    out = {
        "metric": f"flops/hour, {len(cfgs)} batched flops x {args.flop_iters} DCFR iterations ({args.storage} arenas)",
        "value": round(per_hour, 2), "unit": "flops/hour", "n_gpus": 1, "steps": args.flop_iters, "warmup": 0,
        "ms_per_step": round(dt * 1e3 / max(1, args.flop_iters), 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "compressed": "i16/u16 arenas, f32 math", "half": "f16 arenas, f32 math"}[args.storage],
        "data": "synthetic",
        "config": {"workload": f"batched_flops: {len(cfgs)} seeded canonical flops, per-flop menus from "
                               "{25,33,50,75,125,a}, 1176-combo Beta(0.7,0.7) ranges",
                   "nodes_total": int(sum(r["nodes"] for r in rep["flops"])), "waves": rep["waves"],
                   "parallelism": "single GPU, engines of a wave in one parent hipGraph",
                   "schedules": [{k: sc[k] for k in ("wave", "show_nodes", "per_level")} for sc in rep.get("schedules", [])]},
        "roofline": None, "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
