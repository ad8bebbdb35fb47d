"""Seeded synthetic spots for the benchmark configurations.

Each generator returns a SpotConfig whose ranges are per-combo tokens
(``AhKh:0.73412354``) printed with the shortest text that round-trips the
f32 weight, so the native range parser (range.rs:28-58) reproduces the
weights bit for bit.

  river_spot     5-card board, 400 combos per player, U(0,1] weights,
                 pot 20 / stack 90, bets {33,75,a}, raise 2.5x, one raise.
  turn_spot      4-card board, all 1128 non-conflicting combos,
                 Beta(0.7,0.7) weights, pot 12 / stack 94,
                 bets {33,75,a} turn and river, raises {2.5x,a}.
  flop_srp_spot  3-card board, OOP 550 / IP 650 combos of 1176,
                 U(0.2,1] weights, pot 5.5 / stack 97.25, flop {33,75},
                 turn {75,a}, river {75,a}, raise 2.5x, one raise per street.
  batched_flops  184 canonical flops (cards.rs:236-259) with per-flop bet
                 menus drawn from {25,33,50,75,125,a} and turn-spot weights.
"""
import numpy as np

from .spot import SpotConfig, StreetSizing, TreeConfig, canonical_flops, card_str, combo_from_index

ALL_CARDS = list(range(52))


def _board(rng, n):
    return [int(c) for c in rng.choice(52, size=n, replace=False)]


def _free_combos(board):
    bm = 0
    for c in board:
        bm |= 1 << c
    out = []
    for idx in range(1326):
        hi, lo = combo_from_index(idx)
        if not ((1 << hi) | (1 << lo)) & bm:
            out.append(idx)
    return out


def _range_text(idxs, weights):
    toks = []
    for idx, w in zip(idxs, weights):
        hi, lo = combo_from_index(idx)
        w32 = np.float32(w)
        if not (w32 > 0):
            w32 = np.float32(1e-6)
        toks.append(f"{card_str(hi)}{card_str(lo)}:{str(w32)}")
    return ",".join(toks)


def _sizing(bet, raise_):
    return StreetSizing.of(bet, raise_)


def _beta(rng, n):
    w = rng.beta(0.7, 0.7, size=n).astype(np.float32)
    return np.maximum(w, np.float32(1e-6))


def river_spot(seed=1):
    rng = np.random.default_rng(seed)
    board = _board(rng, 5)
    free = _free_combos(board)
    rngs = []
    for _ in range(2):
        idxs = sorted(int(x) for x in rng.choice(free, size=400, replace=False))
        w = (1.0 - rng.random(400)).astype(np.float32)
        rngs.append(_range_text(idxs, w))
    s = [_sizing("", ""), _sizing("", ""), _sizing("33 75 a", "2.5x")]
    tree = TreeConfig(starting_pot=20.0, effective_stack=90.0, oop=s, ip=[StreetSizing(**vars(x)) for x in s],
                      max_raises=1)
    return SpotConfig("".join(card_str(c) for c in board), rngs[0], rngs[1], tree)


def turn_spot(seed=2):
    rng = np.random.default_rng(seed)
    board = _board(rng, 4)
    free = _free_combos(board)
    rngs = [_range_text(free, _beta(rng, len(free))) for _ in range(2)]
    s = [_sizing("", ""), _sizing("33 75 a", "2.5x a"), _sizing("33 75 a", "2.5x a")]
    tree = TreeConfig(starting_pot=12.0, effective_stack=94.0, oop=s, ip=[StreetSizing(**vars(x)) for x in s])
    return SpotConfig("".join(card_str(c) for c in board), rngs[0], rngs[1], tree)


def flop_srp_spot(seed=3):
    rng = np.random.default_rng(seed)
    board = _board(rng, 3)
    free = _free_combos(board)
    rngs = []
    for n in (550, 650):
        idxs = sorted(int(x) for x in rng.choice(free, size=n, replace=False))
        w = (1.0 - 0.8 * rng.random(n)).astype(np.float32)
        rngs.append(_range_text(idxs, w))
    s = [_sizing("33 75", "2.5x"), _sizing("75 a", "2.5x"), _sizing("75 a", "2.5x")]
    tree = TreeConfig(starting_pot=5.5, effective_stack=97.25, oop=s, ip=[StreetSizing(**vars(x)) for x in s],
                      max_raises=1)
    return SpotConfig("".join(card_str(c) for c in board), rngs[0], rngs[1], tree)


def batched_flops(n=184, seed=4, max_raises=0):
    """Seeded per-flop menus over canonical flops; yields SpotConfigs."""
    rng = np.random.default_rng(seed)
    menu = ["25", "33", "50", "75", "125", "a"]
    out = []
    for board, _w in canonical_flops(n):
        cards = [board[i:i + 2] for i in (0, 2, 4)]
        from .spot import parse_cards
        free = _free_combos(parse_cards("".join(cards)))
        rngs = [_range_text(free, _beta(rng, len(free))) for _ in range(2)]
        k = int(rng.integers(1, 3))
        sizes = " ".join(sorted(rng.choice(menu, size=k, replace=False), key=menu.index))
        s = [_sizing(sizes, "2.5x"), _sizing(sizes, "2.5x"), _sizing(sizes, "2.5x")]
        tree = TreeConfig(starting_pot=5.5, effective_stack=97.25, oop=s, ip=[StreetSizing(**vars(x)) for x in s],
                          max_raises=max_raises)
        out.append(SpotConfig(board, rngs[0], rngs[1], tree))
    return out


CONFIGS = {"river": river_spot, "turn": turn_spot, "flop_srp": flop_srp_spot}
