"""The game value of a solved spot and its gradient, on the CPU oracle.

U_p = sum_i w_p[i] * cfv_p(root)[i] (best_response.rs:264-269) is bilinear in
the two ranges, so the CFR counterfactual values are its exact gradient:
dU_p/dw_p = cfv_p. Without rake the game is zero-sum, U_0 = -U_1, and
dU_0/dw_1 = -cfv_1. Central differences in FP64 check both; a plausible
mistake (dropped same-combo term, a weight applied twice, a sign) breaks one.
"""
import numpy as np
import pytest


def _setup(rake=0.0, iters=40):
    import oracle
    from gtopen_amd import Spot, workloads
    from gtopen_amd.game_value import behavioural_strategy

    cfg = workloads.river_spot()
    cfg.tree.rake_pct = rake
    cfg.tree.rake_cap = 3.0 if rake else 0.0
    spot = Spot(cfg)
    solver = oracle.Oracle(spot)
    solver.iterate(iters)
    _, strat = solver.arenas()
    sigma = behavioural_strategy(spot, strat)
    o64 = oracle.Oracle(spot, precision=64, weights=(spot.weights[0], spot.weights[1]), isomorphism=False)
    o64.set_arenas([np.zeros(spot.data_size[q], dtype=np.float32) for q in range(2)], sigma)
    return spot, sigma, o64


def _fd(o64, spot, sigma, p, q, d, eps):
    w = [spot.weights[0].astype(np.float64), spot.weights[1].astype(np.float64)]
    vals = []
    for s in (1.0, -1.0):
        ww = [w[0].copy(), w[1].copy()]
        ww[q] = ww[q] + s * eps * d
        o64.set_weights(ww[0].astype(np.float32), ww[1].astype(np.float32))
        vals.append(o64.game_value(p, sigma))
    o64.set_weights(spot.weights[0], spot.weights[1])
    # the weights the oracle saw are f32: difference them exactly
    dq = ((w[q] + eps * d).astype(np.float32).astype(np.float64) - (w[q] - eps * d).astype(np.float32)) / 2.0
    return (vals[0] - vals[1]) / (2.0 * eps), dq / eps


def test_cfv_is_the_range_gradient_of_the_game_value(built):
    spot, sigma, o64 = _setup()
    rng = np.random.default_rng(11)
    for p in range(2):
        c_own = o64.root_cfv(p, mode=1)
        c_opp = o64.root_cfv(1 - p, mode=1)
        u = o64.game_value(p, sigma)
        assert u == pytest.approx(float(np.dot(spot.weights[p].astype(np.float64), c_own)), rel=1e-12)
        # own range: dU_p/dw_p = cfv_p
        d = rng.standard_normal(spot.num_hands[p])
        fd, dq = _fd(o64, spot, sigma, p, p, d, 0.05)
        assert fd == pytest.approx(float(np.dot(dq, c_own)), rel=1e-9, abs=1e-9 * np.abs(c_own).sum())
        # zero-sum without rake: dU_p/dw_o = -cfv_o
        d = rng.standard_normal(spot.num_hands[1 - p])
        fd, dq = _fd(o64, spot, sigma, p, 1 - p, d, 0.05)
        assert fd == pytest.approx(-float(np.dot(dq, c_opp)), rel=1e-9, abs=1e-9 * np.abs(c_opp).sum())
    u0, u1 = o64.game_value(0, sigma), o64.game_value(1, sigma)
    assert u0 + u1 == pytest.approx(0.0, abs=1e-9 * abs(u0))


def test_raked_game_is_not_zero_sum(built):
    spot, sigma, o64 = _setup(rake=0.05)
    u0, u1 = o64.game_value(0, sigma), o64.game_value(1, sigma)
    # rake leaves the table: the two values sum to minus the expected rake
    assert u0 + u1 < -1e-6 * abs(u0)


def test_game_value_is_linear_in_one_nodes_strategy(built):
    """Every path crosses a node at most once: U is affine along a
    normalisation-preserving change of one node's strategy."""
    from gtopen_amd import _lib as L

    spot, sigma, o64 = _setup()
    act = np.nonzero(spot.kind == L.KIND_ACTION)[0]
    rng = np.random.default_rng(5)
    for n in act[:4]:
        q = int(spot.player[n])
        na, nh, off = int(spot.num_children[n]), spot.num_hands[q], int(spot.data_offset[n])
        blk = sigma[q][off:off + na * nh].reshape(na, nh).astype(np.float64)
        r = rng.uniform(-1, 1, size=(na, nh))
        d = blk * (r - (blk * r).sum(axis=0))
        vals = []
        for t in (-0.2, 0.0, 0.2):
            s2 = [sigma[0].copy(), sigma[1].copy()]
            s2[q][off:off + na * nh] = (blk + t * d).astype(np.float32).reshape(-1)
            vals.append(o64.game_value(0, s2))
        assert vals[2] - vals[1] == pytest.approx(vals[1] - vals[0], rel=1e-5, abs=1e-7 * abs(vals[1]))
