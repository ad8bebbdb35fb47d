"""Spot configuration and construction (mirrors crates/solver/src/game.rs:21-133
and tree.rs:14-102).

``SpotConfig``/``TreeConfig``/``StreetSizing`` serialise to the reference's
serde JSON layout, so a config file written for the reference's
``solve-cli`` loads unchanged. ``Spot`` builds the tree and river tables in
the native host library and exposes them as numpy views.
"""
import ctypes
import json
from dataclasses import dataclass, field
from typing import List

import numpy as np

from . import _lib as L


def parse_sizes(text):
    """'33 75' / '50, a' / '2.5x' -> serde bet-size list (tree.rs:21-57)."""
    out = []
    for tok in text.replace(",", " ").replace(";", " ").split():
        t = tok.strip().lower()
        if t in ("a", "allin", "all-in"):
            out.append("AllIn")
        elif t.endswith("x"):
            m = float(t[:-1])
            if not (m == m and m > 1.0 and m != float("inf")):
                raise ValueError(f"raise multiple must be a finite number > 1: {tok!r}")
            out.append({"PrevMult": m})
        else:
            p = float(t)
            if not (p == p and p > 0.0 and p != float("inf")):
                raise ValueError(f"bet size must be a positive finite number: {tok!r}")
            out.append({"PotPct": p})
    return out


@dataclass
class StreetSizing:
    bet: list = field(default_factory=list)
    raise_: list = field(default_factory=list)
    donk: list = field(default_factory=list)

    @staticmethod
    def of(bet="", raise_="", donk=""):
        return StreetSizing(parse_sizes(bet), parse_sizes(raise_), parse_sizes(donk))

    def to_json(self):
        return {"bet": self.bet, "raise": self.raise_, "donk": self.donk}


@dataclass
class TreeConfig:
    starting_pot: float = 60.0
    effective_stack: float = 970.0
    rake_pct: float = 0.0
    rake_cap: float = 0.0
    oop: List[StreetSizing] = field(default_factory=lambda: [StreetSizing() for _ in range(3)])
    ip: List[StreetSizing] = field(default_factory=lambda: [StreetSizing() for _ in range(3)])
    allin_threshold: float = 0.85
    add_allin: bool = False
    max_raises: int = 10

    def to_json(self):
        return {
            "starting_pot": self.starting_pot,
            "effective_stack": self.effective_stack,
            "rake_pct": self.rake_pct,
            "rake_cap": self.rake_cap,
            "oop": [s.to_json() for s in self.oop],
            "ip": [s.to_json() for s in self.ip],
            "allin_threshold": self.allin_threshold,
            "add_allin": self.add_allin,
            "max_raises": self.max_raises,
        }


@dataclass
class SpotConfig:
    board: str
    range_oop: str
    range_ip: str
    tree: TreeConfig = field(default_factory=TreeConfig)

    def to_json(self):
        return json.dumps(
            {"board": self.board, "range_oop": self.range_oop, "range_ip": self.range_ip, "tree": self.tree.to_json()}
        )


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).view(dtype)


class Spot:
    """A built spot: tree, hands and river tables (game.rs:82-100).

    ``Spot(config_json_or_SpotConfig, strict=True, max_nodes=0)``.
    """

    def __init__(self, config, strict=True, max_nodes=0):
        text = config.to_json() if isinstance(config, SpotConfig) else (
            config if isinstance(config, str) else json.dumps(config)
        )
        self.config_json = text
        h = ctypes.c_void_p()
        L.check(
            L.lib().gto_spot_build(text.encode(), L.STRICT if strict else L.LENIENT_LOAD, max_nodes, ctypes.byref(h)),
            "spot build",
        )
        self._h = h
        self.view_ptr = L.lib().gto_spot_view_of(h)
        v = self.view_ptr.contents
        self.view = v
        n = v.num_nodes
        self.num_nodes = n
        self.kind = _arr(v.node_kind, n, np.uint8)
        self.player = _arr(v.node_player, n, np.uint8)
        self.street = _arr(v.node_street, n, np.uint8)
        self.num_children = _arr(v.node_num_children, n, np.uint8)
        self.children_start = _arr(v.node_children_start, n, np.uint32)
        self.data_offset = _arr(v.node_data_offset, n, np.uint64)
        self.children = _arr(v.children, v.num_children_total, np.uint32)
        self.data_size = (int(v.data_size[0]), int(v.data_size[1]))
        self.num_hands = (int(v.num_hands[0]), int(v.num_hands[1]))
        self.hand_c1 = [_arr(v.hand_c1[p], self.num_hands[p], np.uint8) for p in range(2)]
        self.hand_c2 = [_arr(v.hand_c2[p], self.num_hands[p], np.uint8) for p in range(2)]
        self.weights = [_arr(v.weights[p], self.num_hands[p], np.float32) for p in range(2)]
        self.same_combo = [_arr(v.same_combo[p], self.num_hands[p], np.uint32) for p in range(2)]
        self.put = _arr(v.node_put, 2 * n, np.float64).reshape(n, 2)
        self.board = [int(v.board[k]) for k in range(int(v.board_len))]
        self.starting_pot = float(v.starting_pot)
        self.root_street = int(v.root_street)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                L.lib().gto_spot_free(h)
            except Exception:
                pass
            self._h = None

    @property
    def num_action_nodes(self):
        return int(L.lib().gto_spot_num_action_nodes(self._h))

    def arena_bytes(self, storage=L.STORAGE_F32):
        return int(L.lib().gto_spot_arena_bytes(self._h, storage))

    @property
    def num_suit_perms(self):
        return int(L.lib().gto_spot_num_suit_perms(self._h))

    @property
    def suit_perms(self):
        """Suit symmetries of board and ranges, identity first (game.rs:262-317):
        symmetry k maps suit s to suit_perms[k][s]."""
        v = self.view
        return [tuple(int(v.suit_perms[4 * k + s]) for s in range(4)) for k in range(int(v.num_suit_perms))]

    def child(self, node, a):
        return int(self.children[int(self.children_start[node]) + a])

    def hand_index(self, p, combo):
        """Index of a combo like 'AcKc' in player p's hand list."""
        cs = parse_cards(combo)
        for i in range(self.num_hands[p]):
            if {int(self.hand_c1[p][i]), int(self.hand_c2[p][i])} == set(cs):
                return i
        raise KeyError(f"combo {combo} not in player {p}'s range")


RANKS = "23456789TJQKA"
SUITS = "cdhs"


def card_str(c):
    return RANKS[c >> 2] + SUITS[c & 3]


def parse_cards(s):
    s = s.replace(" ", "").replace(",", "")
    return [RANKS.index(s[i].upper()) * 4 + SUITS.index(s[i + 1].lower()) for i in range(0, len(s), 2)]


def combo_index(c1, c2):
    hi, lo = max(c1, c2), min(c1, c2)
    return hi * (hi - 1) // 2 + lo


def combo_from_index(idx):
    hi = 1
    while (hi + 1) * hi // 2 <= idx:
        hi += 1
    return hi, idx - hi * (hi - 1) // 2


def range_parse(text):
    w = (ctypes.c_float * 1326)()
    L.check(L.lib().gto_range_parse(text.encode(), w), "range parse")
    return np.frombuffer(w, dtype=np.float32).copy()


def evaluate7(cards):
    a = (ctypes.c_uint8 * 7)(*cards)
    return int(L.lib().gto_evaluate7(a))


def canonical_flops(n=0):
    cap = 1755
    cards = (ctypes.c_uint8 * (3 * cap))()
    w = (ctypes.c_uint32 * cap)()
    cnt = ctypes.c_uint32()
    L.check(L.lib().gto_canonical_flops(n, cards, w, cap, ctypes.byref(cnt)), "canonical flops")
    return [("".join(card_str(cards[3 * k + i]) for i in range(3)), int(w[k])) for k in range(cnt.value)]
