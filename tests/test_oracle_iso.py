"""Suit isomorphism in the oracle (cfr.rs:568-660, 749-870) and
canonical paths (query.rs:643-700).

The iso engine traverses one chance branch per suit orbit and synthesizes
the others through hand permutations; symmetrize then copies the
representatives' arenas onto their orbit members. The engine without
isomorphism traverses every branch and is pinned by the rest of the suite,
so it is the independent check: on a spot with suit symmetry both must give
the same arenas (exactly over the first iterations, where every sum is the
same sum in another order only when ties appear) and the same
exploitability trajectory to f32 rounding. A wrong hand table, a
representative chosen the wrong way round or a copy in the wrong direction
all give O(1) differences in the first iteration.
"""
import numpy as np
import pytest


def symmetric_turn():
    """Four hearts on board, ranges without suit preference: the symmetries
    are the 6 permutations of spades, diamonds and clubs."""
    from gtopen_amd import SpotConfig, StreetSizing, TreeConfig

    return SpotConfig("Th7h4h2h", "AA,KK,QQ,JJ,AK,AQ,KQ,JTs", "AA,KK,QQ,AK,AQ,KQ,T9s,87s",
                      TreeConfig(starting_pot=20.0, effective_stack=100.0,
                                 oop=[StreetSizing.of("50", "") for _ in range(3)],
                                 ip=[StreetSizing.of("50", "") for _ in range(3)]))


def _flat(arenas):
    return [*arenas[0], *arenas[1]]


def test_symmetry_group_and_hand_tables(built):
    from gtopen_amd import Spot

    spot = Spot(symmetric_turn())
    perms = spot.suit_perms
    assert perms[0] == (0, 1, 2, 3) and len(perms) == 6
    assert all(pm[2] == 2 for pm in perms)  # hearts fixed
    v = spot.view
    for q in range(2):
        nh = spot.num_hands[q]
        for k, pm in enumerate(perms):
            tbl = np.array([v.hand_perm[q][k * nh + i] for i in range(nh)])
            assert sorted(tbl) == list(range(nh))  # a permutation of the hands
            for i in range(nh):
                c1, c2 = int(spot.hand_c1[q][i]), int(spot.hand_c2[q][i])
                j = int(tbl[i])
                img = {(c1 & ~3) | pm[c1 & 3], (c2 & ~3) | pm[c2 & 3]}
                assert img == {int(spot.hand_c1[q][j]), int(spot.hand_c2[q][j])}


@pytest.mark.parametrize("storage", ["f32", "compressed"])
def test_iso_matches_every_branch_traversal(built, storage):
    import oracle
    from gtopen_amd import Spot

    spot = Spot(symmetric_turn())
    a = oracle.Oracle(spot, storage=storage)
    b = oracle.Oracle(spot, storage=storage, isomorphism=False)
    pot = spot.starting_pot
    for it in (1, 2):
        a.iterate(1)
        b.iterate(1)
        for x, y in zip(_flat(a.arenas()), _flat(b.arenas())):
            assert np.allclose(x, y, rtol=0, atol=1e-6 * max(1.0, float(np.abs(y).max())))
    for k in range(4):
        a.iterate(10)
        b.iterate(10)
        assert abs(a.exploitability() - b.exploitability()) <= 1e-4 * pot


def test_iso_average_strategy_of_copied_branches(built):
    """A non-representative action node's average strategy (after
    symmetrize) is its representative's with hands permuted."""
    import oracle
    from gtopen_amd import Spot, _lib as L

    spot = Spot(symmetric_turn())
    o = oracle.Oracle(spot)
    o.iterate(15)
    root_chance = next(n for n in range(spot.num_nodes) if spot.kind[n] == L.KIND_CHANCE)
    cs = int(spot.children_start[root_chance])
    # 3s and its representative 3c (card = 4 * rank + suit, suits c d h s)
    three_c, three_s = 1 * 4 + 0, 1 * 4 + 3
    rep, dst = int(spot.children[cs + three_c]), int(spot.children[cs + three_s])
    assert spot.kind[rep] == L.KIND_ACTION and spot.kind[dst] == L.KIND_ACTION
    # the symmetry symmetrize uses: the first one (in symmetry order) giving
    # the smallest image, as its strictly-less scan picks
    best, k = three_s, 0
    for kk, pm in enumerate(spot.suit_perms):
        img = (three_s & ~3) | pm[three_s & 3]
        if img < best:
            best, k = img, kk
    assert best == three_c
    p = int(spot.player[rep])
    nh, na = spot.num_hands[p], int(spot.num_children[rep])
    tbl = np.array([spot.view.hand_perm[p][k * nh + i] for i in range(nh)])
    sr, sd = o.average_strategy(rep), o.average_strategy(dst)
    for a in range(na):
        assert np.array_equal(sd[a * nh:(a + 1) * nh], sr[a * nh:(a + 1) * nh][tbl])


def test_canonical_path(built):
    from gtopen_amd import Spot
    from gtopen_amd.query import QueryMixin
    from gtopen_amd.spot import parse_cards

    class Q(QueryMixin):
        def __init__(self, spot, iso):
            self.spot, self.isomorphism = spot, iso

    spot = Spot(symmetric_turn())
    q = Q(spot, True)
    # clubs are the smallest suit (card = 4 * rank + suit, suits c d h s):
    # a spade or diamond river maps to clubs
    assert parse_cards("3c")[0] == 4
    assert q.canonical_path([0, 0, "3s"]) == [0, 0, "3c"]
    assert q.canonical_path([0, 0, "3d"]) == [0, 0, "3c"]
    assert q.canonical_path([0, 0, "3c"]) == [0, 0, "3c"]
    assert q.canonical_path([0, 0, "3h"]) == [0, 0, "3h"]  # hearts are fixed by every symmetry
    path, composed = q.canonical_path_with_perm([0, "5s"])
    assert path == [0, "5c"] and composed[3] == 0  # the map sends spades to clubs
    assert q.canonical_path([{"type": "card", "card": "Ad"}]) == [{"type": "card", "card": "Ac"}]
    assert Q(spot, False).canonical_path([0, 0, "3s"]) == [0, 0, "3s"]
