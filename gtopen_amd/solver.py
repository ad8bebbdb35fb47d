"""GpuSolver: the MI355X engine behind the reference's GpuSolver interface
(crates/solver/src/gpu/mod.rs:60-997): new / iterate / exploitability /
sync (download) / update_locks / synchronize / iteration, over the C ABI.
"""
import ctypes
from typing import NamedTuple

import numpy as np

from . import _lib as L
from .query import QueryMixin


def schedule_dict(sc):
    return {"show_nodes": int(sc.show_nodes), "per_level": bool(sc.per_level), "river": bool(sc.river),
            "absorb": bool(sc.absorb)}


class Progress(NamedTuple):
    """Progress (cfr.rs:91-97), plus the seconds spent in iterations."""
    iteration: int
    exploit_chips: float
    exploit_pct_pot: float
    elapsed_secs: float
    gpu_secs: float

    @staticmethod
    def of(p):
        return Progress(int(p.iteration), float(p.exploit_chips), float(p.exploit_pct_pot), float(p.elapsed_secs),
                        float(p.gpu_secs))


class GpuSolver(QueryMixin):
    def __init__(self, spot, storage="f32", algo="dcfr", use_graphs=True, device=0, showdown="exact",
                 isomorphism=True):
        """isomorphism: traverse one chance branch per suit orbit (the
        reference's default use_isomorphism, cfr.rs:84-87, 143); downloads
        and strategies of the other branches are their suit-permuted
        copies, as after the reference's ensure_symmetric."""
        self.spot = spot  # keeps the view's arrays alive
        self.isomorphism = bool(isomorphism)
        self.device = int(device)
        opts = L.HipOptions(
            L.STORAGE[storage],
            L.algo_code(algo),
            1 if use_graphs else 0,
            device,
            1 if isomorphism else 0,
        )
        self.storage = storage
        self.locks, self.lock_labels = {}, {}
        h = ctypes.c_void_p()
        L.check(L.lib().gto_hip_create(spot.view_ptr, ctypes.byref(opts), ctypes.byref(h)), "gpu create")
        self._h = h
        if showdown != "exact":
            self.set_showdown(showdown)

    def set_schedule(self, show_nodes=16, per_level=False, river=False, absorb=True):
        """Launch schedule (results are identical under every schedule):
        river = fused river sweeps (one workgroup per river subtree, its rows
        in LDS) where the engine allows them; for the level sweeps, showdown
        nodes per workgroup (16, 8 or 4), whether each level's terminals
        launch with that level, and whether traverser nodes absorb their
        opponent children (f32 DCFR sweeps without locks)."""
        sc = L.Schedule(show_nodes, int(bool(per_level)), int(bool(river)), int(bool(absorb)))
        L.check(L.lib().gto_hip_set_schedule(self._h, ctypes.byref(sc)), "set schedule")

    def schedule(self):
        sc = L.Schedule()
        L.check(L.lib().gto_hip_get_schedule(self._h, ctypes.byref(sc)), "get schedule")
        return schedule_dict(sc)

    def autotune(self, iters=2):
        """Time every candidate schedule over `iters` iterations (plus one
        warm-up each; these are ordinary iterations of the solve), keep the
        fastest. Returns (schedule, [ms per iteration per candidate])."""
        best = L.Schedule()
        ms = np.zeros(L.SCHEDULE_CANDIDATES, dtype=np.float32)
        L.check(L.lib().gto_hip_autotune(self._h, iters, ctypes.byref(best), ms.ctypes.data_as(L.c_f32p)),
                "autotune")
        return schedule_dict(best), [float(x) for x in ms]

    @property
    def river_entries(self):
        """Arena entries inside river subtrees (the fused river kernel's
        share of reached_entries; 0 when fused sweeps are unavailable)."""
        return int(L.lib().gto_hip_river_entries(self._h))

    def set_showdown(self, mode):
        """'exact' (f64 prefix sums, the CPU engine's arithmetic up to the
        association of f64 sums), 'sequential' ('exact' with the fold and
        showdown f64 sums in the CPU engine's order wherever the order could
        change a bit: every sweep, exploitability included, equals the CPU
        engine's bit for bit; f32 rows only), 'bf16' or 'fp8' (same-board
        showdowns stacked into matrix-core products with bf16 or e4m3 reach,
        f32 accumulation). The precision forms serve the DCFR sweeps;
        exploitability and average-strategy values stay exact."""
        L.check(L.lib().gto_hip_set_showdown(self._h, L.SHOWDOWN[mode]), "set showdown")
        self.showdown = mode

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().gto_hip_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def iteration(self):
        it = ctypes.c_uint32()
        L.check(L.lib().gto_hip_get_iteration(self._h, ctypes.byref(it)), "get iteration")
        return it.value

    @iteration.setter
    def iteration(self, t):
        L.check(L.lib().gto_hip_set_iteration(self._h, t), "set iteration")

    def iterate(self, n=1):
        L.check(L.lib().gto_hip_iterate(self._h, n), "gpu iterate")

    def run(self, max_iterations=1000, target_exploit_pct=0.3, check_every=25, on_progress=None):
        """Solver::run (cfr.rs:194-226; RunOptions defaults of cfr.rs:107-115):
        iterate until max_iterations, the target (% of pot) or request_stop();
        exploitability every check_every iterations and at the last one.
        on_progress(Progress) runs after each check; returning False stops.
        Returns the last check's Progress (iteration, exploit_chips,
        exploit_pct_pot, elapsed_secs, gpu_secs)."""
        opts = L.RunOptions(max_iterations, target_exploit_pct, check_every)
        self._stop = ctypes.c_int(0)
        err = []

        def cb(p, _user):
            if on_progress is None:
                return 1
            try:
                return 0 if on_progress(Progress.of(p.contents)) is False else 1
            except Exception as e:  # a Python error stops the run and is raised after it
                err.append(e)
                return 0

        fn = L.PROGRESS_FN(cb)
        out = L.Progress()
        L.check(L.lib().gto_hip_run(self._h, ctypes.byref(opts), ctypes.byref(self._stop), fn, None,
                                    ctypes.byref(out)), "gpu run")
        if err:
            raise err[0]
        return Progress.of(out)

    def request_stop(self):
        """Stop a run() in progress (from another thread) after its current iteration."""
        if getattr(self, "_stop", None) is not None:
            self._stop.value = 1

    def exploitability(self):
        """Exploitability in chips (best_response.rs:255-287)."""
        e = ctypes.c_double()
        L.check(L.lib().gto_hip_exploitability(self._h, ctypes.byref(e)), "gpu exploitability")
        return e.value

    def root_cfv(self, p, mode=0):
        out = np.zeros(self.spot.num_hands[p], dtype=np.float32)
        L.check(
            L.lib().gto_hip_root_cfv(self._h, p, mode, out.ctypes.data_as(L.c_f32p)), "gpu root cfv"
        )
        return out

    def download_arenas(self):
        """(regrets[2], strat[2]) as f32 numpy arrays (sync_to_cpu)."""
        ds = self.spot.data_size
        r = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        s = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        L.check(
            L.lib().gto_hip_download_arenas(
                self._h, *[a.ctypes.data_as(L.c_f32p) for a in (r[0], r[1], s[0], s[1])]
            ),
            "download arenas",
        )
        return r, s

    def download_compressed(self, p):
        """Raw compressed arenas of player p: (i16 regret quants, u16 strategy
        quants, per-node regret scales, per-node strategy scales)."""
        ds, n = self.spot.data_size[p], self.spot.num_nodes
        rq, sq = np.zeros(ds, np.int16), np.zeros(ds, np.uint16)
        rs, ss = np.zeros(n, np.float32), np.zeros(n, np.float32)
        L.check(L.lib().gto_hip_download_compressed(self._h, p, rq.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                                                    sq.ctypes.data_as(L.c_u16p), rs.ctypes.data_as(L.c_f32p),
                                                    ss.ctypes.data_as(L.c_f32p)), "download compressed")
        return rq, sq, rs, ss

    def upload_arenas(self, regrets, strat):
        """f32 arenas (None for an arena, or for a pair, keeps it)."""
        regrets = regrets if regrets is not None else (None, None)
        strat = strat if strat is not None else (None, None)
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float32)
                for a in (regrets[0], regrets[1], strat[0], strat[1])]
        for k, a in enumerate(arrs):
            if a is not None and a.size != self.spot.data_size[k % 2]:
                raise ValueError(f"arena {k} has {a.size} entries, expected {self.spot.data_size[k % 2]}")
        ptrs = [None if a is None else a.ctypes.data_as(L.c_f32p) for a in arrs]
        L.check(L.lib().gto_hip_upload_arenas(self._h, *ptrs), "upload arenas")

    def update_locks(self, locks):
        """locks: {node_index: sigma (na*nh float array)} (gpu/mod.rs:894-905)."""
        off = np.full(self.spot.num_nodes, -1, dtype=np.int64)
        parts = []
        pos = 0
        for node in sorted(locks):
            sig = np.ascontiguousarray(locks[node], dtype=np.float32)
            off[node] = pos
            parts.append(sig)
            pos += sig.size
        sigma = np.concatenate(parts) if parts else np.zeros(1, dtype=np.float32)
        self.locks = {int(n): np.asarray(locks[n], dtype=np.float32) for n in locks}
        L.check(
            L.lib().gto_hip_update_locks(
                self._h, off.ctypes.data_as(L.c_i64p), sigma.ctypes.data_as(L.c_f32p), pos
            ),
            "update locks",
        )

    def save(self, path, locks=None, labels=None):
        """sync_to_cpu + Solver::save: download (decoding compressed arenas)
        and write a GTOSOLVE2 file. Locks default to the ones last given to
        update_locks."""
        from .save import write_save

        regrets, strat = self.download_arenas()
        locks = self.locks if locks is None else locks
        labels = self.lock_labels if labels is None else labels
        write_save(path, self.spot, self.iteration, regrets, strat, locks=locks, labels=labels)

    def synchronize(self):
        L.check(L.lib().gto_hip_synchronize(self._h), "synchronize")

    def profile(self, n=1):
        prof = L.HipProfile()
        L.check(L.lib().gto_hip_profile_iterations(self._h, n, ctypes.byref(prof)), "profile")
        return {
            L.lib().gto_hip_profile_kind_name(k).decode(): (prof.ms[k], prof.launches[k]) for k in range(L.PROF_KINDS)
        }

    @property
    def reached_entries(self):
        return int(L.lib().gto_hip_reached_entries(self._h))


def vram_info(device=0):
    """(free, total) device memory in MiB (vram_info_mb, gpu/mod.rs:953)."""
    fr, tot = ctypes.c_uint64(), ctypes.c_uint64()
    L.check(L.lib().gto_hip_vram_info(device, ctypes.byref(fr), ctypes.byref(tot)), "vram info")
    return fr.value, tot.value
