"""GTOSOLVE2 save files (crates/solver/src/save.rs) over the C ABI.

``write_save`` is Solver::save (save.rs:94-143) given arenas already on the
host, ``SaveFile`` + ``load_solver`` are Solver::load / load_with_storage
(save.rs:145-227): the spot is rebuilt leniently from the saved config, the
locks are validated against it, the arenas are size-checked and handed to a
GPU engine (compressed engines re-encode them per node on upload).
"""
import ctypes

import numpy as np

from . import _lib as L


def _lock_arrays(locks):
    nodes = sorted(int(n) for n in (locks or {}))
    parts = [np.ascontiguousarray(locks[n], dtype=np.float32).reshape(-1) for n in nodes]
    lens = np.array([p.size for p in parts], dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.uint64) if parts else np.zeros(0, np.uint64)
    sigma = np.concatenate(parts) if parts else np.zeros(1, np.float32)
    return np.array(nodes, dtype=np.uint32), offs, lens, sigma


def validate_locks(spot, locks):
    """validate_locks (save.rs:42-70): raises GtoError with the reference's message."""
    nodes, offs, lens, sigma = _lock_arrays(locks)
    L.check(L.lib().gto_validate_locks(spot._h, len(nodes), nodes.ctypes.data_as(L.c_u32p),
                                       offs.ctypes.data_as(L.c_u64p), lens.ctypes.data_as(L.c_u64p),
                                       sigma.ctypes.data_as(L.c_f32p)), "validate locks")


def write_save(path, spot, iteration, regrets, strat, locks=None, labels=None):
    """Write a GTOSOLVE2 file. regrets/strat: per-player f32 arena arrays;
    locks: {node: sigma}; labels: {node: str}."""
    arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (regrets[0], regrets[1], strat[0], strat[1])]
    for k, a in enumerate(arrs):
        if a.size != spot.data_size[k % 2]:
            raise ValueError(f"arena {k} has {a.size} entries, expected {spot.data_size[k % 2]}")
    nodes, offs, lens, sigma = _lock_arrays(locks)
    lab_nodes = np.array(sorted(int(n) for n in (labels or {})), dtype=np.uint32)
    lab_bytes = [str(labels[int(n)]).encode() for n in lab_nodes]
    lab_arr = (ctypes.c_char_p * max(1, len(lab_bytes)))(*lab_bytes)
    L.check(L.lib().gto_save_write(str(path).encode(), spot._h, int(iteration),
                                   *[a.ctypes.data_as(L.c_f32p) for a in arrs],
                                   len(nodes), nodes.ctypes.data_as(L.c_u32p), offs.ctypes.data_as(L.c_u64p),
                                   lens.ctypes.data_as(L.c_u64p), sigma.ctypes.data_as(L.c_f32p),
                                   len(lab_nodes), lab_nodes.ctypes.data_as(L.c_u32p), lab_arr), "save")


class SaveFile:
    """An opened save: header (config, iteration, locks, labels); arenas on demand."""

    def __init__(self, path):
        h = ctypes.c_void_p()
        L.check(L.lib().gto_save_open(str(path).encode(), ctypes.byref(h)), "load")
        self._h = h
        lib = L.lib()
        self.config_json = lib.gto_save_config_json(h).decode()
        self.iteration = int(lib.gto_save_iteration(h))
        self.locks = {}
        for k in range(lib.gto_save_num_locks(h)):
            node, ptr, n = ctypes.c_uint32(), L.c_f32p(), ctypes.c_uint64()
            L.check(lib.gto_save_lock(h, k, ctypes.byref(node), ctypes.byref(ptr), ctypes.byref(n)), "save lock")
            self.locks[int(node.value)] = np.ctypeslib.as_array(ptr, shape=(int(n.value),)).copy() if n.value else \
                np.zeros(0, np.float32)
        self.labels = {}
        for k in range(lib.gto_save_num_labels(h)):
            node, txt = ctypes.c_uint32(), ctypes.c_char_p()
            L.check(lib.gto_save_label(h, k, ctypes.byref(node), ctypes.byref(txt)), "save label")
            self.labels[int(node.value)] = txt.value.decode()

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            L.lib().gto_save_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def spot(self, max_nodes=0):
        """The spot rebuilt leniently from the saved config (save.rs:162-165)."""
        from .spot import Spot

        return Spot(self.config_json, strict=False, max_nodes=max_nodes)

    def validate_locks(self, spot):
        L.check(L.lib().gto_save_validate_locks(self._h, spot._h), "load")

    def read_arenas(self, spot):
        ds = spot.data_size
        r = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        s = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        L.check(L.lib().gto_save_read_arenas(self._h, spot._h, *[a.ctypes.data_as(L.c_f32p)
                                                                  for a in (r[0], r[1], s[0], s[1])]), "load")
        return r, s


def load_solver(path, storage="f32", algo="dcfr", device=0):
    """Solver::load_with_storage for the GPU engine: (spot, GpuSolver, SaveFile)."""
    from .solver import GpuSolver

    sf = SaveFile(path)
    spot = sf.spot()
    sf.validate_locks(spot)
    regrets, strat = sf.read_arenas(spot)
    gpu = GpuSolver(spot, storage=storage, algo=algo, device=device)
    gpu.upload_arenas(regrets, strat)
    gpu.iteration = sf.iteration
    if sf.locks:
        gpu.update_locks(sf.locks)
    gpu.locks = dict(sf.locks)
    gpu.lock_labels = dict(sf.labels)
    return spot, gpu, sf
