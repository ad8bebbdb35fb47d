"""Reports: many flops solved to a target, the role of the reference
server's report worker (crates/server/src/main.rs:1583-1772, report_solve
1774-1825, write_report 1827-1853), redesigned around the GPU batch.

Flops are grouped into waves that fit the device memory (the engines' own
VRAM estimates against free memory); a wave's engines iterate together,
each engine's iteration graph on its own stream (gtopen_amd.batch). Every
``check_every`` iterations each unfinished flop's exploitability is taken;
a flop that reaches the target (or the iteration cap) is recorded and leaves
the batch, so uneven trees finish at their own pace. Rows carry the
reference's fields (board, weight, iterations, exploit_pct, seconds, per
player EV, root action frequencies, IP's response to a root check); the
equity columns of the reference (its equity enumeration) are not produced.
The report file is rewritten through a temporary file every 8 flops and at
the end, with "complete" false until every flop is in.
"""
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import _lib as L
from .batch import Batch
from .solver import GpuSolver, vram_info
from .spot import Spot, SpotConfig, canonical_flops


def flop_configs(template, n_flops):
    """canonical_flops_subset(n) (cards.rs:236-259) boards on a template
    SpotConfig: [(SpotConfig, weight)]."""
    out = []
    for board, weight in canonical_flops(n_flops):
        t = template.tree
        out.append((SpotConfig(board, template.range_oop, template.range_ip, t), weight))
    return out


def _root_stats(gpu, spot):
    """Per-player reach x valid weighted EV at the root, the root actor's
    reach-weighted action frequencies, and IP's frequencies after a root
    check (main.rs:1540-1581, 1714-1729)."""
    players = []
    for p in range(2):
        cur = gpu.node_values(p, 1, [0])[0].astype(np.float64)
        reach = spot.weights[p].astype(np.float64)
        valid = gpu.valid_opp_sum(p, spot.weights[1 - p])
        ok = valid > 1e-9
        w = reach * valid
        ev = np.where(ok, cur / np.where(ok, valid, 1.0) + spot.put[0][p], 0.0)
        wt = float(w[ok].sum())
        players.append({"ev": float((ev[ok] * w[ok]).sum() / wt) if wt > 1e-12 else 0.0})

    def freqs(path):
        v = gpu.node_view(path)
        if v["player"] is None:
            return None
        hs = v["players"][v["player"]]["hands"]
        tot = sum(h["reach"] for h in hs)
        na = len(hs[0]["strategy"]) if hs else 0
        f = [sum(h["strategy"][a] * h["reach"] for h in hs) / tot if tot > 1e-12 else 0.0 for a in range(na)]
        return {"actor": v["player"], "freqs": f}

    root = freqs([])
    vs_check = None
    if spot.kind[0] == L.KIND_ACTION and int(spot.num_children[0]) > 0:
        # the reference's action list puts check first when the root can check
        c = spot.child(0, 0)
        if spot.kind[c] == L.KIND_ACTION and int(spot.player[c]) == 1:
            vs_check = freqs([{"type": "action", "index": 0}])
    return players, root, vs_check


def _write(path, meta, rows, complete):
    if not path:
        return
    out = dict(meta, created=int(time.time()), complete=complete, flops=rows)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(out, f)
    os.replace(tmp, path)


def run_report(flops, target=0.35, max_iterations=600, check_every=20, path=None, storage="f32",
               mem_fraction=0.8, device=0, max_batch=64, name="report", progress=None, autotune=0):
    """Solve [(SpotConfig, weight)] to ``target`` % pot; returns the report
    dict with ``flops_per_hour``. autotune > 0: each wave's batch first times
    every launch schedule over that many iterations (plus a warm-up each;
    these count as iterations of the solves) and keeps the fastest; the
    chosen schedules are listed under ``schedules``."""
    t0 = time.time()
    meta = {"name": name, "target_pct": target, "max_iterations": max_iterations, "storage": storage}
    rows = []
    _write(path, meta, rows, False)
    free_mb, _total_mb = vram_info(device)
    budget = free_mb * (1 << 20) * mem_fraction
    st = L.STORAGE[storage]
    k = 0
    waves = 0
    schedules = []
    tune_cost = L.SCHEDULE_CANDIDATES * (autotune + 1)
    needs = []
    pending = None
    # spots are built ahead on host threads (the C++ builder runs outside the
    # interpreter lock), at most `window` beyond the flop being placed, so
    # building overlaps the device's iterations of the current wave
    pool = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    window = max_batch
    built = {}

    def spot_of(j):
        for t in range(j, min(len(flops), j + window)):
            if t not in built:
                built[t] = pool.submit(Spot, flops[t][0])
        return built.pop(j).result()

    try:
        while k < len(flops):
            # waves of even size: the memory budget sets how many waves the
            # remaining flops need (from the mean estimate so far), each wave
            # takes its share, and the budget is never exceeded (at least one
            # flop per wave); a short last wave would leave the device idle
            if pending is None:
                pending = (flops[k][0], flops[k][1], spot_of(k))
                needs.append(int(L.lib().gto_hip_estimate_vram(pending[2].view_ptr, st)))
            remaining = len(flops) - k
            mean = sum(needs) / len(needs)
            n_waves = max(1, int(np.ceil(mean * remaining / budget)), int(np.ceil(remaining / max_batch)))
            cap = int(np.ceil(remaining / n_waves))
            wave, used = [], 0
            while k < len(flops) and len(wave) < cap:
                if pending is None:
                    pending = (flops[k][0], flops[k][1], spot_of(k))
                    needs.append(int(L.lib().gto_hip_estimate_vram(pending[2].view_ptr, st)))
                need = needs[-1]
                if wave and used + need > budget:
                    break
                wave.append(pending)
                pending = None
                used += need
                k += 1
            waves += 1
            tw = time.time()
            # the wave's engines are created on the same threads (plan building
            # and uploads of independent spots)
            gpus = list(pool.map(lambda w: GpuSolver(w[2], storage=storage, device=device), wave))
            active = list(range(len(wave)))
            batch = None
            tuned = False
            while active:
                if batch is None:
                    batch = Batch([gpus[i] for i in active])
                    if autotune > 0 and not tuned and gpus[active[0]].iteration + tune_cost <= max_iterations:
                        sched, ms = batch.autotune(autotune)
                        schedules.append(dict(sched, wave=waves, ms_per_iteration=ms))
                    tuned = True
                it = gpus[active[0]].iteration
                step = min(check_every - it % check_every, max_iterations - it)
                batch.iterate(step)
                it += step
                done = []
                # the engines' checks run concurrently (each engine has its own
                # stream; a small tree's pass alone leaves most of the device idle)
                pcts = list(pool.map(lambda i: gpus[i].exploitability() / wave[i][2].starting_pot * 100.0, active))
                fin = [i for i, pct in zip(active, pcts) if pct <= target or it >= max_iterations]
                stats = dict(zip(fin, pool.map(lambda i: _root_stats(gpus[i], wave[i][2]), fin)))
                for i, pct in zip(active, pcts):
                    cfg, weight, spot = wave[i]
                    if i in stats:
                        players, root, vs_check = stats[i]
                        rows.append({"board": cfg.board, "weight": weight, "iterations": it, "exploit_pct": pct,
                                     "seconds": time.time() - tw, "nodes": spot.num_nodes,
                                     "arena_entries": gpus[i].reached_entries, "players": players,
                                     "root": root, "vs_check": vs_check})
                        done.append(i)
                        if progress:
                            progress(rows[-1])
                        if len(rows) % 8 == 0:
                            _write(path, meta, rows, False)
                if done:
                    batch.close()
                    batch = None
                    for i in done:
                        gpus[i].close()
                    active = [i for i in active if i not in done]
            del gpus
    finally:
        # queued spot builds are dropped if a wave fails
        pool.shutdown(wait=True, cancel_futures=True)
    elapsed = time.time() - t0
    meta.update(seconds=elapsed, waves=waves, flops_per_hour=len(rows) / elapsed * 3600.0 if elapsed > 0 else 0.0,
                schedules=schedules)
    _write(path, meta, rows, True)
    return dict(meta, complete=True, flops=rows)
