"""Host core (native spot builder) against the reference's own tests and facts:
cards.rs:123-215, evaluator.rs:208-392, range.rs:305-377, the bench_spot
tree facts of bench/new_f32_dcfr.log:1 and new_i16_dcfr.log:1, and the
build errors of tests/build_validation.rs."""
import itertools
import json

import numpy as np
import pytest

from conftest import golden


# ---- cards --------------------------------------------------------------------------------

def test_combo_index_roundtrip(built):
    from gtopen_amd import combo_from_index, combo_index, lib

    seen = set()
    for c1 in range(52):
        for c2 in range(c1):
            idx = lib().gto_combo_index(c1, c2)
            assert idx == combo_index(c1, c2) < 1326
            assert combo_from_index(idx) == (c1, c2)
            seen.add(idx)
    assert len(seen) == 1326


def test_canonical_flops(built):
    from gtopen_amd import canonical_flops

    allf = canonical_flops(0)
    assert len(allf) == 1755  # realization.rs:8-14
    assert sum(w for _, w in allf) == 22100
    for n in (47, 95, 184):  # cards.rs:177-185: unit weights at the UI presets
        sub = canonical_flops(n)
        assert len(sub) == n and all(w == 1 for _, w in sub)
    for n in (1, 500, 923, 924, 1000, 1754):  # cards.rs:158-175
        sub = canonical_flops(n)
        assert len({b for b, _ in sub}) == len(sub) and sum(w for _, w in sub) == n


# ---- evaluator: an independent best-of-21 five-card evaluator ---------------------------

def _eval5(cards):
    ranks = sorted((c >> 2 for c in cards), reverse=True)
    suits = [c & 3 for c in cards]
    flush = len(set(suits)) == 1
    counts = {}
    for r in ranks:
        counts[r] = counts.get(r, 0) + 1
    groups = sorted(((c, r) for r, c in counts.items()), reverse=True)
    uniq = sorted(counts, reverse=True)
    straight = None
    if len(uniq) == 5:
        if uniq[0] - uniq[4] == 4:
            straight = uniq[0]
        elif uniq == [12, 3, 2, 1, 0]:
            straight = 3
    if flush and straight is not None:
        return (8, [straight])
    if groups[0][0] == 4:
        return (7, [groups[0][1], groups[1][1]])
    if groups[0][0] == 3 and groups[1][0] == 2:
        return (6, [groups[0][1], groups[1][1]])
    if flush:
        return (5, ranks)
    if straight is not None:
        return (4, [straight])
    if groups[0][0] == 3:
        return (3, [g[1] for g in groups])
    if groups[0][0] == 2 and groups[1][0] == 2:
        return (2, [g[1] for g in groups])
    if groups[0][0] == 2:
        return (1, [g[1] for g in groups])
    return (0, ranks)


def _eval7_ref(cards):
    return max(_eval5(c) for c in itertools.combinations(cards, 5))


class XorShift:  # evaluator.rs:300-311
    def __init__(self, s):
        self.s = s

    def next(self):
        x = self.s
        x ^= (x << 13) & 0xFFFFFFFFFFFFFFFF
        x ^= x >> 7
        x ^= (x << 17) & 0xFFFFFFFFFFFFFFFF
        self.s = x
        return x


def test_evaluator_matches_best_of_21(built):
    from gtopen_amd import evaluate7

    rng = XorShift(0x9E3779B97F4A7C15)  # the reference test's seed (evaluator.rs:342)
    for _ in range(3000):
        deck = list(range(52))
        for i in range(14):
            j = i + rng.next() % (52 - i)
            deck[i], deck[j] = deck[j], deck[i]
        h1, h2 = deck[:7], deck[7:14]
        f1, f2 = evaluate7(h1), evaluate7(h2)
        r1, r2 = _eval7_ref(h1), _eval7_ref(h2)
        assert f1 >> 20 == r1[0] and f2 >> 20 == r2[0]
        assert (f1 > f2) - (f1 < f2) == (r1 > r2) - (r1 < r2)


def test_evaluator_known_hands(built):
    from gtopen_amd import evaluate7, parse_cards

    ev = lambda s: evaluate7(parse_cards(s))  # noqa: E731
    assert ev("AsKsQsJsTs2c2d") > ev("AcAdAhAs2c2d2h")
    assert ev("As2s3s4s5s9c9d") >> 20 == 8  # wheel straight flush
    assert ev("KcKdKh3c3d2h2s") > ev("KcKdKh2c2d4h5s")
    assert ev("6c5d4h3s2cKdQd") > ev("Ac5d4h3s2cKdQd")  # wheel is the lowest straight
    assert ev("KcKdKh2c2d4h4s") > ev("KcKdKh3c3d3h9s")
    assert ev("AsKsQsJsTs2c3d") == ev("AsKsQsJsTs4c5d")


# ---- range parser (range.rs:305-377) ------------------------------------------------------

def test_range_counts_and_weights(built):
    from gtopen_amd import GtoError, combo_index, range_parse

    cnt = lambda s: float(range_parse(s).sum())  # noqa: E731
    assert cnt("AA") == 6 and cnt("AKs") == 4 and cnt("AKo") == 12 and cnt("AK") == 16
    assert cnt("AhKh") == 1 and cnt("TT+") == 30 and cnt("ATs+") == 16 and cnt("99-66") == 24
    assert cnt("A5s-A2s") == 16 and cnt("T9s-87s") == 12 and cnt("AA,KK,QQ") == 18
    a1, a2 = 12 * 4 + 0, 12 * 4 + 1
    assert abs(range_parse("AA:0.5,KK")[combo_index(a1, a2)] - 0.5) < 1e-6
    assert abs(range_parse("AA:50")[combo_index(a1, a2)] - 0.5) < 1e-6
    assert abs(range_parse("AA,AA:0.25")[combo_index(a1, a2)] - 0.25) < 1e-6
    assert abs(range_parse("AKs:1.5")[combo_index(12 * 4, 11 * 4)] - 0.015) < 1e-6
    for bad in ("XX", "AAs", "AKs:abc"):
        with pytest.raises(GtoError):
            range_parse(bad)


# ---- tree builder ---------------------------------------------------------------------------

def test_bench_spot_tree_facts(built):
    """bench/new_f32_dcfr.log:1 and new_i16_dcfr.log:1 of the reference."""
    from gtopen_amd import Spot

    s = Spot(open(golden("bench_spot.json")).read())
    assert s.num_nodes == 1245971
    assert s.num_action_nodes == 500748
    assert s.num_hands == (257, 319)
    assert f"{s.arena_bytes(0) / 1e6:.1f}" == "2494.8"
    assert f"{s.arena_bytes(1) / 1e6:.1f}" == "1267.3"


def test_benchmark_config_sizes(built):
    from gtopen_amd import Spot, workloads

    sizes = {"river": 33, "turn": 28509, "flop_srp": 1445723}
    for name, fn in workloads.CONFIGS.items():
        s = Spot(fn())
        assert s.num_nodes == sizes[name], name
    t = Spot(workloads.turn_spot())
    assert t.num_hands == (1128, 1128)


def _cfg(**tree):
    from gtopen_amd import SpotConfig, StreetSizing, TreeConfig

    base = dict(starting_pot=60.0, effective_stack=200.0,
                oop=[StreetSizing.of("50", "60") for _ in range(3)],
                ip=[StreetSizing.of("50", "60") for _ in range(3)])
    base.update(tree)
    return SpotConfig("Th9h2c", "QQ,JJ,TT,99,AhKh,AsKs,87s", "AA,KK,AQs,JTs,T9s", TreeConfig(**base))


def test_build_validation_errors(built):
    """Errors of tests/build_validation.rs and game.rs:141-160, 211-226."""
    from gtopen_amd import GtoError, Spot, SpotConfig, StreetSizing

    for bad in (dict(rake_pct=5.0), dict(rake_pct=-0.1), dict(rake_cap=-1.0), dict(starting_pot=0.0),
                dict(effective_stack=float("nan"))):
        with pytest.raises(GtoError):
            Spot(_cfg(**bad))
    with pytest.raises(GtoError):  # raise-only size in a bet list (strict build)
        Spot(_cfg(oop=[StreetSizing(bet=[{"PrevMult": 2.5}]) for _ in range(3)]))
    # lenient load silently drops it (tree.rs:213-243)
    Spot(_cfg(oop=[StreetSizing(bet=[{"PrevMult": 2.5}]) for _ in range(3)]), strict=False)
    with pytest.raises(GtoError):  # node budget
        Spot(_cfg(), max_nodes=100)
    with pytest.raises(GtoError):  # mutually impossible ranges
        c = _cfg()
        Spot(SpotConfig("Th9h2c", "AsKs", "AsQs", c.tree))
    with pytest.raises(GtoError):
        Spot(SpotConfig("Th9h", "AA", "KK", _cfg().tree))


def test_json_roundtrip_reference_layout(built):
    """A solve-cli JSON config (serde layout) builds the same tree as the
    dataclass form."""
    from gtopen_amd import Spot

    cfg = _cfg()
    a = Spot(cfg)
    b = Spot(json.loads(cfg.to_json()))
    assert a.num_nodes == b.num_nodes and a.data_size == b.data_size
    assert np.array_equal(a.data_offset, b.data_offset)
