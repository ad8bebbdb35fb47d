"""Node queries and node locks over the GPU engine: the solver crate's query
layer (crates/solver/src/query.rs) for a solve held in VRAM.

Paths are lists of steps in the reference's serde form,
``{"type": "action", "index": i}`` / ``{"type": "card", "card": "4h"}``
(an int or a card string is accepted as shorthand). Lock modes are
``{"kind": "freeze"}``, ``{"kind": "range", "freqs": [...]}`` and
``{"kind": "hands", "edits": [{"combo": "AcAd", "freqs": [...]}]}``.

Strategies come from the device strategy arena (gto_hip_solved_strategy),
per-node best-response and average-strategy values from one device sweep
each (gto_hip_node_values); walking, lock construction and view assembly
run on the host, as in the reference. Under suit isomorphism (the
engine's default) CFR traverses one chance branch per suit orbit:
canonical_path maps a browsed path onto the traversed branch
(query.rs:643-700) and views of other branches are read through the
composed suit map; with isomorphism off, or a spot without symmetries, it
is the identity.
"""
import numpy as np

from . import _lib as L
from .spot import card_str, parse_cards

SENTINEL = 0xFFFFFFFF
KIND_NAMES = {L.KIND_ACTION: "action", L.KIND_CHANCE: "chance", L.KIND_FOLD: "terminal_fold",
              L.KIND_SHOWDOWN: "terminal_showdown"}


class QueryError(ValueError):
    pass


def _step(s):
    if isinstance(s, dict):
        if s.get("type") == "action":
            return ("action", int(s["index"]))
        if s.get("type") == "card":
            return ("card", str(s["card"]))
        raise QueryError(f"bad path step {s!r}")
    if isinstance(s, (int, np.integer)):
        return ("action", int(s))
    if isinstance(s, str):
        return ("card", s)
    raise QueryError(f"bad path step {s!r}")


def normalize_freqs(freqs, na):
    """A per-action frequency vector scaled to sum 1 (query.rs:40-62): the
    sum is taken in f64; negative, non-finite, all-zero or absurdly large
    vectors are refused."""
    f = np.asarray(freqs, dtype=np.float32).reshape(-1)
    if f.size != na:
        raise QueryError(f"expected {na} action frequencies, got {f.size}")
    if np.any(~np.isfinite(f)) or np.any(f < 0):
        raise QueryError("frequencies must be finite and >= 0")
    total = float(np.sum(f.astype(np.float64)))
    if not np.isfinite(total) or total > 1e9:
        raise QueryError(f"frequencies sum to {total:e} — not a frequency vector (use values on a 0..1 scale)")
    if total <= 1e-9:
        raise QueryError("at least one frequency must be positive")
    return (f.astype(np.float64) / total).astype(np.float32)


def rake_to_target(sigma, reach, target):
    """Scale each action of a node's strategy by one multiplier per action so
    the reach-weighted aggregate frequencies hit ``target`` (query.rs:64-150):
    iterative proportional fitting, up to 400 rounds, stopping once every
    targeted action is within 1e-4; multipliers are renormalised to sum to
    the action count each round. Actions with a zero target are driven out;
    hands holding no mass on any targeted action take the target itself.
    sigma: (na, nh) f32, returns the new (na, nh) f32 strategy."""
    sig = np.array(sigma, dtype=np.float32, copy=True)
    na, nh = sig.shape
    tgt = np.asarray(target, dtype=np.float32)
    on = tgt > 1e-9
    m = on.astype(np.float64)
    mass = np.zeros(nh)
    for a in np.nonzero(on)[0]:
        mass += sig[a].astype(np.float64)
    dead = mass <= 1e-12
    sig[:, dead] = tgt[:, None]
    w = np.asarray(reach, dtype=np.float32).astype(np.float64)
    s64 = sig.astype(np.float64)
    for _ in range(400):
        ms = m[:, None] * s64
        denom = ms.sum(axis=0)
        use = (w > 0.0) & (denom > 1e-12)
        total = float(w[use].sum())
        if total <= 1e-12:
            break
        agg = (ms[:, use] * (w[use] / denom[use])).sum(axis=1)
        f = agg / total
        err = float(np.max(np.abs(f - tgt)[on])) if on.any() else 0.0
        up = on & (f > 1e-9)
        m[up] *= tgt[up] / f[up]
        msum = m.sum()
        if msum > 1e-12:
            m *= na / msum
        if err < 1e-4:
            break
    v32 = (m[:, None] * s64).astype(np.float32)
    tot = (m[:, None] * s64).sum(axis=0)
    out = np.where(tot > 1e-12, (v32.astype(np.float64) / np.where(tot > 1e-12, tot, 1.0)), tgt[:, None])
    return out.astype(np.float32)


class Walk:
    def __init__(self, node, reach, dealt):
        self.node_idx = node
        self.reach = reach
        self.dealt = dealt


class QueryMixin:
    """Query and lock methods of the GPU solver (query.rs:296-961)."""

    # --- strategies --------------------------------------------------------
    def solved_strategy(self, node):
        """Average strategy from the arena, ignoring locks (cfr.rs:679-726)."""
        sp = self.spot
        if sp.kind[node] != L.KIND_ACTION:
            raise QueryError(f"node {node} is not an action node")
        p = int(sp.player[node])
        out = np.zeros(int(sp.num_children[node]) * sp.num_hands[p], dtype=np.float32)
        L.check(L.lib().gto_hip_solved_strategy(self._h, int(node), out.ctypes.data_as(L.c_f32p)), "solved strategy")
        return out

    def average_strategy(self, node):
        """The lock where one is set, else the solved strategy (cfr.rs:668-674)."""
        if node in self.locks:
            return np.asarray(self.locks[node], dtype=np.float32).copy()
        return self.solved_strategy(node)

    def node_values(self, p, mode, nodes):
        """(len(nodes), nh_p) value rows: mode 0 best response, 1 average
        strategy, 2 current strategies (with the engine's showdown form)."""
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        out = np.zeros((nodes.size, self.spot.num_hands[p]), dtype=np.float32)
        L.check(L.lib().gto_hip_node_values(self._h, p, mode, nodes.size, nodes.ctypes.data_as(L.c_u32p),
                                            out.ctypes.data_as(L.c_f32p)), "node values")
        return out

    # --- paths -------------------------------------------------------------
    def canonical_path(self, path):
        """Card steps mapped to their suit-orbit representatives, so
        operations land on the branch CFR traverses (query.rs:643-647)."""
        return self.canonical_path_with_perm(path)[0]

    def canonical_path_with_perm(self, path):
        """Canonical path plus the composed suit map from the browsed branch
        to the canonical one (query.rs:649-700): each card is first moved by
        the map so far, then replaced by its smallest image under the
        symmetries fixing the canonical cards dealt before it (the same
        strictly-less scan as the engine's symmetrize)."""
        composed = [0, 1, 2, 3]
        perms = self.spot.suit_perms
        if not getattr(self, "isomorphism", False) or len(perms) < 2:
            return list(path), composed

        def permute(c, pm):
            return (c & ~3) | pm[c & 3]

        dealt, out = [], []
        for st in path:
            kind, val = _step(st)
            cs = parse_cards(val) if kind == "card" else []
            if kind != "card" or len(cs) != 1:
                out.append(st)
                continue
            cm = permute(cs[0], composed)
            rep, k_rep = cm, None
            for k, pm in enumerate(perms):
                if any(permute(d, pm) != d for d in dealt):
                    continue
                pc = permute(cm, pm)
                if pc < rep:
                    rep, k_rep = pc, k
            if k_rep is not None:
                pm = perms[k_rep]
                composed = [pm[s] for s in composed]
            dealt.append(rep)
            out.append({"type": "card", "card": card_str(rep)} if isinstance(st, dict) else card_str(rep))
        return out, composed

    def walk_path(self, path):
        """Follow a path from the root, multiplying the acting player's reach
        by the average strategy and zeroing hands blocked by dealt cards
        (query.rs:304-350)."""
        sp = self.spot
        node = 0
        reach = [sp.weights[0].astype(np.float32).copy(), sp.weights[1].astype(np.float32).copy()]
        dealt = []
        masks = [(1 << sp.hand_c1[q].astype(np.uint64)) | (1 << sp.hand_c2[q].astype(np.uint64)) for q in range(2)]
        for st in path:
            kind, val = _step(st)
            k = int(sp.kind[node])
            if k == L.KIND_ACTION and kind == "action":
                na = int(sp.num_children[node])
                if not 0 <= val < na:
                    raise QueryError(f"action index {val} out of range")
                p = int(sp.player[node])
                nh = sp.num_hands[p]
                sig = self.average_strategy(node)
                reach[p] = (reach[p] * sig[val * nh:(val + 1) * nh]).astype(np.float32)
                node = sp.child(node, val)
            elif k == L.KIND_CHANCE and kind == "card":
                cs = parse_cards(val)
                if len(cs) != 1:
                    raise QueryError(f"bad card {val!r}")
                c = cs[0]
                child = int(sp.children[int(sp.children_start[node]) + c])
                if child == SENTINEL or c in dealt:
                    raise QueryError(f"card {val} not available here")
                for q in range(2):
                    reach[q][(masks[q] >> np.uint64(c)) & np.uint64(1) == 1] = 0.0
                dealt.append(c)
                node = child
            else:
                raise QueryError("path step does not match node type")
        return Walk(node, reach, dealt)

    # --- locks -------------------------------------------------------------
    def lock_node(self, path, mode, label=""):
        """Lock the node at ``path`` (query.rs:702-776). Every mode starts from
        a stable base, so re-applying never compounds: freeze = the solved
        strategy; range = the solved strategy raked to the target aggregate
        frequencies under the actor's reach; hands = the current strategy
        (lock or solved) with the named combos set exactly."""
        sp = self.spot
        walk = self.walk_path(self.canonical_path(path))
        n = walk.node_idx
        if sp.kind[n] != L.KIND_ACTION:
            raise QueryError("only action nodes can be locked")
        na, p = int(sp.num_children[n]), int(sp.player[n])
        nh = sp.num_hands[p]
        kind = mode.get("kind") if isinstance(mode, dict) else str(mode)
        if kind == "freeze":
            sig = self.solved_strategy(n)
        elif kind == "range":
            target = normalize_freqs(mode["freqs"], na)
            sig = rake_to_target(self.solved_strategy(n).reshape(na, nh), walk.reach[p], target).reshape(-1)
        elif kind == "hands":
            sig = self.average_strategy(n).reshape(na, nh)
            for e in mode["edits"]:
                target = normalize_freqs(e["freqs"], na)
                cs = parse_cards(e["combo"])
                if len(cs) != 2:
                    raise QueryError(f"expected a 2-card combo, got {e['combo']!r}")
                hit = np.nonzero(((sp.hand_c1[p] == cs[0]) & (sp.hand_c2[p] == cs[1]))
                                 | ((sp.hand_c1[p] == cs[1]) & (sp.hand_c2[p] == cs[0])))[0]
                if hit.size == 0:
                    raise QueryError(f"{e['combo']} is not in {'OOP' if p == 0 else 'IP'}'s range here")
                sig[:, hit[0]] = target
            sig = sig.reshape(-1)
        else:
            raise QueryError(f"unknown lock mode {mode!r}")
        locks = dict(self.locks)
        locks[n] = np.asarray(sig, dtype=np.float32)
        self.update_locks(locks)
        self.lock_labels[n] = str(label)
        return n

    def unlock_node(self, path):
        walk = self.walk_path(self.canonical_path(path))
        self.lock_labels.pop(walk.node_idx, None)
        if walk.node_idx not in self.locks:
            return False
        locks = dict(self.locks)
        del locks[walk.node_idx]
        self.update_locks(locks)
        return True

    def clear_locks(self):
        self.update_locks({})
        self.lock_labels = {}

    def list_locks(self):
        return sorted(self.lock_labels.values())

    # --- views -------------------------------------------------------------
    def _combo(self, q, i):
        return card_str(int(self.spot.hand_c1[q][i])) + card_str(int(self.spot.hand_c2[q][i]))

    def node_view(self, path):
        """Per-hand reach and (for the actor) average strategy at a node; the
        strategy part of the reference's node_view (query.rs:490-640)."""
        sp = self.spot
        walk = self.walk_path(path)
        n = walk.node_idx
        k = int(sp.kind[n])
        actor = int(sp.player[n]) if k == L.KIND_ACTION else None
        sig = self.average_strategy(n) if actor is not None else None
        players = []
        for q in range(2):
            nh = sp.num_hands[q]
            na = int(sp.num_children[n]) if actor == q else 0
            hands = []
            for i in range(nh):
                h = {"combo": self._combo(q, i), "weight": float(sp.weights[q][i]), "reach": float(walk.reach[q][i])}
                h["strategy"] = [float(sig[a * nh + i]) for a in range(na)] if actor == q else None
                hands.append(h)
            players.append({"hands": hands})
        return {"node": n, "node_type": KIND_NAMES[k], "player": actor, "locked": n in self.locks,
                "pot": float(sp.put[n].sum()), "players": players}

    def valid_opp_sum(self, p, reach_o):
        """Opponent reach compatible with each of p's hands, f64
        (query.rs:465-488)."""
        sp, o = self.spot, 1 - p
        r = np.asarray(reach_o, dtype=np.float32).astype(np.float64)
        t = 0.0
        s = np.zeros(52)
        for j in range(sp.num_hands[o]):
            t += r[j]
            s[sp.hand_c1[o][j]] += r[j]
            s[sp.hand_c2[o][j]] += r[j]
        same = sp.same_combo[p].astype(np.int64)
        same_r = np.where(same != SENTINEL, r[np.where(same != SENTINEL, same, 0)], 0.0)
        return t - s[sp.hand_c1[p]] - s[sp.hand_c2[p]] + same_r

    def exploit_view(self, path, p):
        """Best-response view for exploiter p at a node (query.rs:800-961):
        per hand the best-response and current-strategy EVs, the gain, and
        at the exploiter's own nodes the best-response action mix (the lock
        where the node is locked, else the argmax set within
        max(|best|*1e-5, 1e-5)) and per-action EVs. Averages weight hands by
        reach x compatible opponent reach."""
        if p not in (0, 1):
            raise QueryError("exploiter must be 0 (OOP) or 1 (IP)")
        sp = self.spot
        walk = self.walk_path(self.canonical_path(path))
        n = walk.node_idx
        k = int(sp.kind[n])
        nh = sp.num_hands[p]
        valid = self.valid_opp_sum(p, walk.reach[1 - p])
        is_actor = k == L.KIND_ACTION and int(sp.player[n]) == p
        na = int(sp.num_children[n]) if k == L.KIND_ACTION else 0
        lock = self.locks.get(n)
        if is_actor:
            kids = [sp.child(n, a) for a in range(na)]
            acfv = self.node_values(p, 0, kids)
            if lock is not None:
                lk = np.asarray(lock, np.float32).reshape(na, nh)
                br = np.zeros(nh, np.float32)
                for a in range(na):
                    br = (br + lk[a] * acfv[a]).astype(np.float32)
            else:
                br = acfv.max(axis=0)
        else:
            br = self.node_values(p, 0, [n])[0]
        cur = self.node_values(p, 1, [n])[0]
        put = float(sp.put[n][p])

        def ev(c, i):
            return float(np.float32(float(c) / valid[i] + put)) if valid[i] > 1e-9 else None

        hands = []
        wsum = bsum = csum = 0.0
        for i in range(nh):
            be, ce = ev(br[i], i), ev(cur[i], i)
            h = {"combo": self._combo(p, i), "weight": float(sp.weights[p][i]), "reach": float(walk.reach[p][i]),
                 "valid": float(np.float32(valid[i])), "br_ev": be, "cur_ev": ce,
                 "gain": (float(np.float32(be - ce)) if be is not None and ce is not None else None),
                 "br_strategy": None, "evs": None}
            if is_actor:
                if lock is not None:
                    h["br_strategy"] = [float(lk[a, i]) for a in range(na)]
                else:
                    m = float(acfv[:, i].max())
                    eps = max(abs(m) * 1e-5, 1e-5)
                    win = [m - float(acfv[a, i]) <= eps for a in range(na)]
                    cnt = max(1, sum(win))
                    h["br_strategy"] = [1.0 / cnt if w else 0.0 for w in win]
                h["evs"] = [ev(acfv[a, i], i) for a in range(na)]
            if be is not None and ce is not None:
                w = float(np.float32(walk.reach[p][i])) * float(np.float32(valid[i]))
                if w > 0.0:
                    wsum += w
                    bsum += w * be
                    csum += w * ce
            hands.append(h)
        avg = (bsum / wsum, csum / wsum, (bsum - csum) / wsum) if wsum > 1e-12 else (None, None, None)
        return {"node_type": KIND_NAMES[k], "exploiter": p, "player": int(sp.player[n]) if k == L.KIND_ACTION else None,
                "locked": n in self.locks, "hands": hands, "avg_br_ev": avg[0], "avg_cur_ev": avg[1],
                "avg_gain": avg[2]}
