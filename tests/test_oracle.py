"""Pins of the CPU oracle: the reference's golden trajectories, closed
forms, brute force on small inputs and the reference's quantizer tests."""
import numpy as np
import pytest

from conftest import golden

# bench/new_f32_dcfr.log:2-11 and bench/new_i16_dcfr.log:2-11 (exploitability
# in chips after 10, 20, ... 100 iterations on bench_spot.json).
GOLDEN_F32 = [33.7904, 9.3381, 4.8733, 3.3154, 2.3677, 1.7728, 1.3505, 1.0533, 0.8316, 0.6747]
GOLDEN_I16 = [33.5076, 9.2809, 4.8721, 3.3099, 2.3737, 1.7740, 1.3562, 1.0590, 0.8350, 0.6787]


def _bench_spot():
    from gtopen_amd import Spot

    return Spot(open(golden("bench_spot.json")).read())


def _trajectory(storage, n_checks):
    import oracle

    spot = _bench_spot()
    o = oracle.Oracle(spot, storage=storage)
    out = []
    for _ in range(n_checks):
        o.iterate(10)
        out.append(round(o.exploitability(), 4))
    return out


def test_golden_f32_trajectory_first_points(built):
    assert _trajectory("f32", 2) == GOLDEN_F32[:2]


def test_golden_compressed_trajectory_first_point(built):
    assert _trajectory("compressed", 1) == GOLDEN_I16[:1]


@pytest.mark.slow
def test_golden_f32_trajectory_full(built):
    assert _trajectory("f32", 10) == GOLDEN_F32


@pytest.mark.slow
def test_golden_compressed_trajectory_full(built):
    assert _trajectory("compressed", 10) == GOLDEN_I16


# ---- terminal math against brute force -------------------------------------------------------

def _small_river_spot(seed=5, n=60):
    from gtopen_amd import Spot, SpotConfig, StreetSizing, TreeConfig, card_str, combo_from_index

    rng = np.random.default_rng(seed)
    board = [int(c) for c in rng.choice(52, 5, replace=False)]
    bm = sum(1 << c for c in board)
    free = [i for i in range(1326) if not (sum(1 << c for c in combo_from_index(i)) & bm)]
    rngs = []
    for _ in range(2):
        idx = sorted(rng.choice(free, n, replace=False))
        rngs.append(",".join(f"{card_str(combo_from_index(i)[0])}{card_str(combo_from_index(i)[1])}:"
                             f"{np.float32(1 - rng.random())}" for i in idx))
    s = [StreetSizing(), StreetSizing(), StreetSizing.of("50", "")]
    tree = TreeConfig(starting_pot=10.0, effective_stack=50.0, oop=s, ip=s, max_raises=1)
    return Spot(SpotConfig("".join(card_str(c) for c in board), rngs[0], rngs[1], tree)), board


def test_fold_and_showdown_match_pairwise_brute_force(built):
    import oracle
    from gtopen_amd import evaluate7

    spot, board = _small_river_spot()
    rng = np.random.default_rng(9)
    for p in range(2):
        o = 1 - p
        reach = rng.random(spot.num_hands[o]).astype(np.float32)
        reach[rng.random(reach.size) < 0.2] = 0.0
        mine = [(int(spot.hand_c1[p][i]), int(spot.hand_c2[p][i])) for i in range(spot.num_hands[p])]
        opp = [(int(spot.hand_c1[o][j]), int(spot.hand_c2[o][j])) for j in range(spot.num_hands[o])]
        st_m = [evaluate7(board + list(h)) for h in mine]
        st_o = [evaluate7(board + list(h)) for h in opp]
        fold = oracle.fold_cfv(spot, p, reach, 7.5)
        show = oracle.showdown_cfv(spot, p, 0, reach, 4.0, -5.0, -0.25)
        for i, hm in enumerate(mine):
            valid = win = lose = 0.0
            for j, ho in enumerate(opp):
                if set(hm) & set(ho):
                    continue
                valid += reach[j]
                if st_m[i] > st_o[j]:
                    win += reach[j]
                elif st_m[i] < st_o[j]:
                    lose += reach[j]
            tie = valid - win - lose
            assert abs(fold[i] - 7.5 * valid) <= 1e-5 * (1 + abs(7.5 * valid))
            ref = 4.0 * win - 5.0 * lose - 0.25 * tie
            assert abs(show[i] - ref) <= 1e-5 * (1 + abs(ref)), (p, i, show[i], ref)


# ---- quantizers (store.rs:221-262) ----------------------------------------------------------

def test_quantizer_roundtrip_and_edge_blocks(built):
    import ctypes

    import oracle
    from gtopen_amd import _lib as G

    L = oracle.lib()
    src = (np.sin(np.arange(1000, dtype=np.float32) * np.float32(0.7)) * np.float32(1500.0)
           - np.float32(200.0)).astype(np.float32)
    q = np.zeros(1000, dtype=np.int16)
    scale = L.oracle_encode_i16(src.ctypes.data_as(G.c_f32p), 1000, q.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)))
    back = q.astype(np.float32) * np.float32(scale / np.float32(32767.0))
    assert np.all(np.abs(src - back) <= scale / 32767.0 * 0.5 + 1e-6 * 1500)
    q2 = np.zeros(1000, dtype=np.int16)
    scale2 = L.oracle_encode_i16(back.ctypes.data_as(G.c_f32p), 1000, q2.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)))
    assert scale2 == scale and np.array_equal(q, q2)
    z = np.zeros(4, dtype=np.float32)
    qz = np.full(4, 7, dtype=np.int16)
    assert L.oracle_encode_i16(z.ctypes.data_as(G.c_f32p), 4, qz.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))) == 0
    assert np.all(qz == 0)
    bad = np.array([np.nan, np.inf], dtype=np.float32)
    qb = np.full(2, 7, dtype=np.int16)
    assert L.oracle_encode_i16(bad.ctypes.data_as(G.c_f32p), 2, qb.ctypes.data_as(ctypes.POINTER(ctypes.c_int16))) == 0
    assert np.all(qb == 0)
    us = (np.abs(np.cos(np.arange(1000, dtype=np.float32) * np.float32(0.37))) * np.float32(9.5)).astype(np.float32)
    uq = np.zeros(1000, dtype=np.uint16)
    us_scale = L.oracle_encode_u16(us.ctypes.data_as(G.c_f32p), 1000, uq.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))
    uback = uq.astype(np.float32) * np.float32(us_scale / np.float32(65535.0))
    assert np.all(np.abs(us - uback) <= us_scale / 65535.0 * 0.5 + 1e-6)


# ---- closed forms and invariants -------------------------------------------------------------

def _clairvoyance():
    from gtopen_amd import Spot, SpotConfig, StreetSizing, TreeConfig

    tree = TreeConfig(starting_pot=100.0, effective_stack=100.0,
                      oop=[StreetSizing(), StreetSizing(), StreetSizing.of("100", "")],
                      ip=[StreetSizing(), StreetSizing(), StreetSizing()])
    return Spot(SpotConfig("QcJc9c3d2d", "AcKc,8h7h", "QdQh", tree))


def test_clairvoyance_closed_form(built):
    """convergence.rs:31-105: OOP bets all nuts and half its air, IP calls half."""
    import oracle

    spot = _clairvoyance()
    o = oracle.Oracle(spot)
    o.iterate(3000)
    assert o.exploitability() / 100.0 * 100.0 < 0.10
    sig = o.average_strategy(0)
    nh = spot.num_hands[0]
    nuts, air = spot.hand_index(0, "AcKc"), spot.hand_index(0, "8h7h")
    assert sig[nh + nuts] > 0.97
    assert abs(sig[nh + air] - 0.5) < 0.05
    bet_child = spot.child(0, 1)
    sig_ip = o.average_strategy(bet_child)
    assert abs(sig_ip[spot.num_hands[1] + spot.hand_index(1, "QdQh")] - 0.5) < 0.07
    # zero-sum without rake: OOP's average-strategy value is minus IP's
    v0 = o.root_cfv(0, 1)
    v1 = o.root_cfv(1, 1)
    u0 = float(np.dot(spot.weights[0], v0)) / o.pair_weight_sum()
    u1 = float(np.dot(spot.weights[1], v1)) / o.pair_weight_sum()
    assert abs(u0 + u1) < 1e-4


def test_zero_sum_and_fp64_agreement(built):
    """Without rake the average-strategy values sum to zero; the f32 oracle
    tracks the FP64 oracle closely over 30 iterations."""
    import oracle
    from gtopen_amd import Spot, workloads

    spot = Spot(workloads.river_spot())
    o32, o64 = oracle.Oracle(spot, 32), oracle.Oracle(spot, 64)
    o32.iterate(30)
    o64.iterate(30)
    d = o32.pair_weight_sum()
    u = [float(np.dot(spot.weights[p].astype(np.float64), o64.root_cfv(p, 1))) / d for p in range(2)]
    assert abs(u[0] + u[1]) < 1e-9 * spot.starting_pot
    assert abs(o32.exploitability() - o64.exploitability()) / spot.starting_pot < 1e-4


def test_thread_count_does_not_change_results(built):
    """Chance-branch results merge in card order (cfr.rs:637-663)."""
    import oracle
    from gtopen_amd import Spot, workloads

    spot = Spot(workloads.turn_spot())
    old = oracle.get_threads()
    try:
        res = []
        for th in (1, 4):
            oracle.set_threads(th)
            o = oracle.Oracle(spot)
            o.iterate(2)
            res.append(o.arenas())
        for k in range(2):
            for p in range(2):
                assert np.array_equal(res[0][k][p], res[1][k][p])
    finally:
        oracle.set_threads(old)
