"""Batches of GPU engines iterated together (gto_batch_*): each batch
iteration launches every engine's captured iteration graph on the engine's
own stream, engines interleaved, so independent spots (the flops of a
report) share the device."""
import ctypes

from . import _lib as L


class Batch:
    def __init__(self, solvers):
        self.solvers = list(solvers)  # keeps the engines alive
        if not self.solvers:
            raise ValueError("a batch needs at least one solver")
        arr = (ctypes.c_void_p * len(self.solvers))(*[s._h.value for s in self.solvers])
        h = ctypes.c_void_p()
        L.check(L.lib().gto_batch_create(arr, len(self.solvers), ctypes.byref(h)), "batch create")
        self._h = h

    def iterate(self, n=1):
        L.check(L.lib().gto_batch_iterate(self._h, n), "batch iterate")

    def autotune(self, iters=2):
        """Every candidate launch schedule applied to all engines and timed
        over `iters` batch iterations (plus one warm-up each; ordinary
        iterations of every solve); the fastest stays applied. Returns
        (schedule, [ms per batch iteration per candidate])."""
        best = L.Schedule()
        ms = (ctypes.c_float * L.SCHEDULE_CANDIDATES)()
        L.check(L.lib().gto_batch_autotune(self._h, iters, ctypes.byref(best), ms), "batch autotune")
        from .solver import schedule_dict

        return schedule_dict(best), [float(x) for x in ms]

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().gto_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
