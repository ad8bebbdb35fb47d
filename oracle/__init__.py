"""CPU oracle (TEST INFRASTRUCTURE): ctypes wrapper of oracle/_build/liboracle.so.

Imported only by tests/, by __graft_entry__.smoke() and by bench.py's
cpu_baseline leg, as the checker and the timed CPU baseline. It restates the
reference CPU engine (crates/solver/src/cfr.rs, best_response.rs,
game.rs:470-577, store.rs:153-215); see oracle.h for the function list.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def build():
    subprocess.run(["make", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        from gtopen_amd import _lib as G

        h = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sp = ctypes.POINTER(G.SpotView)
        f32p, f64p, i64p = G.c_f32p, G.c_f64p, G.c_i64p
        sig = [
            ("oracle_create", vp, [sp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            ("oracle_destroy", None, [vp]),
            ("oracle_set_threads", None, [ctypes.c_int]),
            ("oracle_get_threads", ctypes.c_int, []),
            ("oracle_iterate", None, [vp, ctypes.c_uint32]),
            ("oracle_iteration", ctypes.c_uint32, [vp]),
            ("oracle_set_iteration", None, [vp, ctypes.c_uint32]),
            ("oracle_exploitability", ctypes.c_double, [vp]),
            ("oracle_pair_weight_sum", ctypes.c_double, [vp]),
            ("oracle_root_cfv", None, [vp, ctypes.c_int, ctypes.c_int, f64p]),
            ("oracle_get_arenas", None, [vp, f32p, f32p, f32p, f32p]),
            ("oracle_set_arenas", None, [vp, f32p, f32p, f32p, f32p]),
            ("oracle_get_arenas_f64", None, [vp, f64p, f64p, f64p, f64p]),
            ("oracle_set_locks", None, [vp, i64p, f32p, ctypes.c_uint64]),
            ("oracle_average_strategy", None, [vp, ctypes.c_uint32, f32p]),
            ("oracle_set_isomorphism", None, [vp, ctypes.c_int]),
            ("oracle_fold_cfv", None, [sp, ctypes.c_int, f32p, ctypes.c_float, f32p]),
            ("oracle_showdown_cfv", None, [sp, ctypes.c_int, ctypes.c_uint32, f32p, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float, f32p]),
            ("oracle_encode_i16", ctypes.c_float, [f32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int16)]),
            ("oracle_encode_u16", ctypes.c_float, [f32p, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint16)]),
        ]
        for name, res, args in sig:
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def set_threads(n):
    lib().oracle_set_threads(int(n))


def get_threads():
    return int(lib().oracle_get_threads())


class Oracle:
    """CPU DCFR engine over a built gtopen_amd.Spot.

    precision 32 = the reference's f32 arithmetic; 64 = FP64 oracle.
    storage 'f32' or 'compressed' (i16/u16 arenas, precision 32 only).
    """

    def __init__(self, spot, precision=32, storage="f32", algo="dcfr", weights=None, isomorphism=True):
        """weights=(w0, w1): per-hand range weights replacing the spot's (the
        engine reads them through a private copy of the spot's view, so
        set_weights can change them later)."""
        self.spot = spot
        view_ptr = spot.view_ptr
        if weights is not None:
            from gtopen_amd import _lib as G

            self._w = [np.array(weights[q], dtype=np.float32).reshape(spot.num_hands[q]).copy() for q in range(2)]
            self._view = G.SpotView.from_buffer_copy(spot.view_ptr.contents)
            for q in range(2):
                self._view.weights[q] = self._w[q].ctypes.data_as(G.c_f32p)
            view_ptr = ctypes.pointer(self._view)
        key = algo.lower().replace("-", "").replace("_", "").replace("+", "")
        codes = {"dcfr": 0, "cfr": 1, "cfrplus": 1, "pcfr": 2, "pcfrplus": 2}  # Algorithm::parse, cfr.rs:26-34
        if key not in codes:
            raise ValueError(f"unknown algorithm {algo!r} (use dcfr, cfr+ or pcfr+)")
        self._h = lib().oracle_create(view_ptr, precision, 0 if storage == "f32" else 1, codes[key])
        if not isomorphism:
            lib().oracle_set_isomorphism(self._h, 0)

    def set_weights(self, w0, w1):
        if not hasattr(self, "_w"):
            raise RuntimeError("create the oracle with weights=... to change them")
        self._w[0][:] = w0
        self._w[1][:] = w1

    def game_value(self, p, sigma):
        """U_p = sum_i w_p[i] cfv_p(root)[i] with the strategy arenas set to
        sigma (per-player arena arrays, already normalised), average-strategy
        traversal (best_response.rs:128-223, 264-269). Overwrites the
        arenas."""
        zeros = [np.zeros(self.spot.data_size[q], dtype=np.float32) for q in range(2)]
        self.set_arenas(zeros, sigma)
        c = self.root_cfv(p, mode=1)
        w = self._w[p] if hasattr(self, "_w") else self.spot.weights[p]
        return float(np.dot(w.astype(np.float64), c))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().oracle_destroy(h)
            self._h = None

    def iterate(self, n=1):
        lib().oracle_iterate(self._h, n)

    @property
    def iteration(self):
        return int(lib().oracle_iteration(self._h))

    @iteration.setter
    def iteration(self, t):
        lib().oracle_set_iteration(self._h, t)

    def exploitability(self):
        return float(lib().oracle_exploitability(self._h))

    def pair_weight_sum(self):
        return float(lib().oracle_pair_weight_sum(self._h))

    def root_cfv(self, p, mode=0):
        out = np.zeros(self.spot.num_hands[p], dtype=np.float64)
        lib().oracle_root_cfv(self._h, p, mode, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def arenas(self):
        from gtopen_amd import _lib as G

        ds = self.spot.data_size
        r = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        s = [np.zeros(ds[p], dtype=np.float32) for p in range(2)]
        lib().oracle_get_arenas(self._h, *[a.ctypes.data_as(G.c_f32p) for a in (r[0], r[1], s[0], s[1])])
        return r, s

    def arenas_f64(self):
        ds = self.spot.data_size
        r = [np.zeros(ds[p], dtype=np.float64) for p in range(2)]
        s = [np.zeros(ds[p], dtype=np.float64) for p in range(2)]
        dp = ctypes.POINTER(ctypes.c_double)
        lib().oracle_get_arenas_f64(self._h, *[a.ctypes.data_as(dp) for a in (r[0], r[1], s[0], s[1])])
        return r, s

    def set_arenas(self, regrets, strat):
        from gtopen_amd import _lib as G

        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (regrets[0], regrets[1], strat[0], strat[1])]
        lib().oracle_set_arenas(self._h, *[a.ctypes.data_as(G.c_f32p) for a in arrs])

    def set_locks(self, locks):
        from gtopen_amd import _lib as G

        off = np.full(self.spot.num_nodes, -1, dtype=np.int64)
        parts, pos = [], 0
        for node in sorted(locks):
            sig = np.ascontiguousarray(locks[node], dtype=np.float32)
            off[node] = pos
            parts.append(sig)
            pos += sig.size
        sigma = np.concatenate(parts) if parts else np.zeros(1, dtype=np.float32)
        self._lock_keep = (off, sigma)
        lib().oracle_set_locks(self._h, off.ctypes.data_as(G.c_i64p), sigma.ctypes.data_as(G.c_f32p), pos)

    def average_strategy(self, node):
        from gtopen_amd import _lib as G

        p = int(self.spot.player[node])
        out = np.zeros(int(self.spot.num_children[node]) * self.spot.num_hands[p], dtype=np.float32)
        lib().oracle_average_strategy(self._h, node, out.ctypes.data_as(G.c_f32p))
        return out


def fold_cfv(spot, p, reach_o, amount):
    from gtopen_amd import _lib as G

    reach_o = np.ascontiguousarray(reach_o, dtype=np.float32)
    out = np.zeros(spot.num_hands[p], dtype=np.float32)
    lib().oracle_fold_cfv(spot.view_ptr, p, reach_o.ctypes.data_as(G.c_f32p), amount, out.ctypes.data_as(G.c_f32p))
    return out


def showdown_cfv(spot, p, key, reach_o, win, lose, tie):
    from gtopen_amd import _lib as G

    reach_o = np.ascontiguousarray(reach_o, dtype=np.float32)
    out = np.zeros(spot.num_hands[p], dtype=np.float32)
    lib().oracle_showdown_cfv(spot.view_ptr, p, key, reach_o.ctypes.data_as(G.c_f32p), win, lose, tie,
                              out.ctypes.data_as(G.c_f32p))
    return out
