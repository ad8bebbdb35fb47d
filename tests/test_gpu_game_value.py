"""GPU game value and its hand-written backward (gto_hip_game_value) against
the FP64 oracle: the value, the CFR counterfactual values (the reference
gradient), FP64 central finite differences for ranges and strategies, and
the autograd wrapper. Tolerances: relative 2e-5 of the quantity's scale
(f32 sweeps with f64 terminal sums against FP64).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RTOL = 2e-5


def _say(msg):
    print(msg, flush=True)


def _solved(cfg, iters):
    import oracle
    from gtopen_amd import GameValue, GpuSolver, Spot
    from gtopen_amd.game_value import behavioural_strategy

    spot = Spot(cfg)
    g = GpuSolver(spot)
    g.iterate(iters)
    _, strat = g.download_arenas()
    g.close()
    sigma = behavioural_strategy(spot, strat)
    gv = GameValue(spot)
    gv.set_strategy(*sigma)
    o64 = oracle.Oracle(spot, precision=64, weights=(spot.weights[0], spot.weights[1]), isomorphism=False)
    o64.set_arenas([np.zeros(spot.data_size[q], dtype=np.float32) for q in range(2)], sigma)
    return spot, sigma, gv, o64


def _close(x, y, scale, what):
    err = float(np.max(np.abs(np.asarray(x, np.float64) - np.asarray(y, np.float64))))
    _say(f"  {what}: max |gpu-fp64| {err:.3e} (scale {scale:.3e})")
    assert err <= RTOL * scale, what


def test_game_value_and_range_gradient_turn_spot(built):
    from gtopen_amd import workloads

    _say("case: turn spot (1128x1128), 100 GPU DCFR iterations, frozen average strategy")
    spot, sigma, gv, o64 = _solved(workloads.turn_spot(), 100)
    for p in range(2):
        o = 1 - p
        u, gw, gs = gv.value_and_grad(p)
        cp, co = o64.root_cfv(p, mode=1), o64.root_cfv(o, mode=1)
        u_ref = o64.game_value(p, sigma)
        wp, wo = spot.weights[p].astype(np.float64), spot.weights[o].astype(np.float64)
        scale_u = float(np.dot(wp, np.abs(cp)))
        _say(f"  player {p}: U gpu {u:.6f} fp64 {u_ref:.6f}")
        assert abs(u - u_ref) <= RTOL * scale_u
        _close(gw[p], cp, np.abs(cp).max(), f"p{p} dU/dw_own vs cfv")
        # no rake: zero-sum, the opponent-range gradient is minus its cfv
        _close(gw[o], -co, np.abs(co).max(), f"p{p} dU/dw_opp vs -cfv_opp")
        # bilinearity: U = w_p . dU/dw_p = w_o . dU/dw_o
        assert float(np.dot(wp, gw[p])) == pytest.approx(u, abs=RTOL * scale_u)
        assert float(np.dot(wo, gw[o])) == pytest.approx(u, abs=RTOL * scale_u)
        # every path crosses the root: sum sigma * dU/dsigma there is U
        root_q = int(spot.player[0])
        n0 = int(spot.num_children[0]) * spot.num_hands[root_q]
        s_root = float(np.dot(sigma[root_q][:n0].astype(np.float64), gs[root_q][:n0].astype(np.float64)))
        assert s_root == pytest.approx(u, abs=RTOL * scale_u)


@pytest.mark.timeout(900)
def test_range_perturbations_against_fp64(built):
    """The differentiable-game-value workload: the turn spot solved for 500
    DCFR iterations (the trajectory's converged end), strategies frozen, 64
    range perturbations N(0, 0.02^2) clipped to [0,1]. At every one, the
    value against the FP64 oracle and both range gradients against the
    oracle's CFR counterfactual values (dU/dw_p = cfv_p, dU/dw_o = -cfv_o
    without rake); FP64 central differences on 4 of them."""
    from gtopen_amd import workloads

    _say("case: turn spot (500 iterations), 64 range perturbations N(0, 0.02^2) clipped to [0,1]; FD on 4 of them")
    spot, sigma, gv, o64 = _solved(workloads.turn_spot(), 500)
    rng = np.random.default_rng(2024)
    w = [spot.weights[q].astype(np.float64) for q in range(2)]
    worst, worst_g = 0.0, 0.0
    for k in range(64):
        wk = [np.clip(w[q] + rng.normal(0.0, 0.02, size=w[q].shape), 0.0, 1.0).astype(np.float32) for q in range(2)]
        gv.set_weights(*wk)
        o64.set_weights(*wk)
        u, gw, _ = gv.value_and_grad(0, grad_strategy=False)
        u_ref = o64.game_value(0, sigma)
        c0 = o64.root_cfv(0, mode=1)
        c1 = o64.root_cfv(1, mode=1)
        scale = float(np.dot(wk[0].astype(np.float64), np.abs(c0)))
        worst = max(worst, abs(u - u_ref) / scale)
        assert abs(u - u_ref) <= RTOL * scale, f"perturbation {k}"
        for g, ref in ((gw[0], c0), (gw[1], -c1)):
            err = float(np.abs(g.astype(np.float64) - ref).max()) / float(np.abs(ref).max())
            worst_g = max(worst_g, err)
            assert err <= RTOL, f"perturbation {k}: gradient vs CFR values {err:.3e}"
        if k % 16 == 0:
            # FP64 central differences along random directions in both ranges
            for q in range(2):
                d = rng.standard_normal(spot.num_hands[q])
                eps = 0.05
                vals = []
                for s in (1.0, -1.0):
                    ww = [wk[0].astype(np.float64), wk[1].astype(np.float64)]
                    ww[q] = ww[q] + s * eps * d
                    o64.set_weights(ww[0].astype(np.float32), ww[1].astype(np.float32))
                    vals.append(o64.game_value(0, sigma))
                o64.set_weights(*wk)
                dq = ((wk[q] + eps * d).astype(np.float32).astype(np.float64)
                      - (wk[q] - eps * d).astype(np.float32)) / (2.0 * eps)
                fd = (vals[0] - vals[1]) / (2.0 * eps)
                lin = float(np.dot(dq, gw[q].astype(np.float64)))
                mag = float(np.dot(np.abs(dq), np.abs(gw[q].astype(np.float64))))
                _say(f"  perturbation {k} range {q}: FD {fd:.6f} grad.d {lin:.6f}")
                assert abs(fd - lin) <= RTOL * mag
    _say(f"  worst relative value error over 64 perturbations: {worst:.3e}; gradient vs CFR values {worst_g:.3e}")


def test_strategy_gradient_against_fp64_differences(built):
    from gtopen_amd import _lib as L
    from gtopen_amd import workloads

    _say("case: turn spot, dU/dsigma on single-node perturbations (U is linear in one node's strategy)")
    spot, sigma, gv, o64 = _solved(workloads.turn_spot(), 100)
    u, _, gs = gv.value_and_grad(0, grad_weights=False)
    rng = np.random.default_rng(7)
    act = np.nonzero(spot.kind == L.KIND_ACTION)[0]
    picks = [0] + [int(x) for x in rng.choice(act[1:], size=11, replace=False)]
    for n in picks:
        q = int(spot.player[n])
        na, nh, off = int(spot.num_children[n]), spot.num_hands[q], int(spot.data_offset[n])
        blk = sigma[q][off:off + na * nh].reshape(na, nh).astype(np.float64)
        r = rng.uniform(-1.0, 1.0, size=(na, nh))
        d = blk * (r - (blk * r).sum(axis=0))  # keeps every hand's sum and sign
        eps = 0.4
        vals, used = [], []
        for s in (1.0, -1.0):
            s2 = [sigma[0].copy(), sigma[1].copy()]
            blk32 = (blk + s * eps * d).astype(np.float32)
            s2[q][off:off + na * nh] = blk32.reshape(-1)
            used.append(blk32.astype(np.float64))
            vals.append(o64.game_value(0, s2))
        fd = (vals[0] - vals[1]) / (2.0 * eps)
        g = gs[q][off:off + na * nh].reshape(na, nh).astype(np.float64)
        lin = float(np.sum((used[0] - used[1]) / (2.0 * eps) * g))
        mag = float(np.sum(np.abs(d) * np.abs(g))) + 1e-30
        _say(f"  node {n} (player {q}, {na} actions): FD {fd:.6f} grad.d {lin:.6f}")
        assert abs(fd - lin) <= 1e-4 * mag + RTOL * abs(u) * 1e-3


def test_raked_cross_gradient_against_fp64(built):
    """With rake the game is not zero-sum: the opponent-range gradient comes
    from the swapped-payoff sweep and is checked by FP64 differences."""
    from gtopen_amd import workloads

    _say("case: river spot with 5% rake (cap 3), opponent-range gradient by FP64 differences")
    cfg = workloads.river_spot()
    cfg.tree.rake_pct = 0.05
    cfg.tree.rake_cap = 3.0
    spot, sigma, gv, o64 = _solved(cfg, 60)
    rng = np.random.default_rng(3)
    for p in range(2):
        o = 1 - p
        u, gw, _ = gv.value_and_grad(p, grad_strategy=False)
        co = o64.root_cfv(o, mode=1)
        assert np.max(np.abs(gw[o] + co)) > 1e-3 * np.abs(co).max()  # rake breaks -cfv
        d = rng.standard_normal(spot.num_hands[o])
        eps = 0.05
        vals = []
        for s in (1.0, -1.0):
            ww = [spot.weights[0].astype(np.float64), spot.weights[1].astype(np.float64)]
            ww[o] = ww[o] + s * eps * d
            o64.set_weights(ww[0].astype(np.float32), ww[1].astype(np.float32))
            vals.append(o64.game_value(p, sigma))
        o64.set_weights(spot.weights[0], spot.weights[1])
        dq = ((spot.weights[o] + eps * d).astype(np.float32).astype(np.float64)
              - (spot.weights[o] - eps * d).astype(np.float32)) / (2.0 * eps)
        fd = (vals[0] - vals[1]) / (2.0 * eps)
        lin = float(np.dot(dq, gw[o].astype(np.float64)))
        mag = float(np.dot(np.abs(dq), np.abs(gw[o].astype(np.float64))))
        _say(f"  player {p}: FD {fd:.6f} grad.d {lin:.6f}")
        assert abs(fd - lin) <= RTOL * mag


def test_autograd_function(built):
    import torch

    from gtopen_amd import game_value, workloads
    from gtopen_amd.game_value import to_combos

    _say("case: river spot, torch.autograd.Function backward")
    spot, sigma, gv, o64 = _solved(workloads.river_spot(), 40)
    dev = "cuda"
    w0 = torch.tensor(spot.weights[0], device=dev, requires_grad=True)
    w1 = torch.tensor(spot.weights[1], device=dev, requires_grad=True)
    s0 = torch.tensor(sigma[0], device=dev, requires_grad=True)
    s1 = torch.tensor(sigma[1], device=dev, requires_grad=True)
    u = game_value(gv, 0, w0, w1, s0, s1)
    (2.0 * u).backward()
    u2, gw, gs = gv.value_and_grad(0)
    assert float(u.detach()) == pytest.approx(u2, rel=1e-12)
    np.testing.assert_allclose(w0.grad.cpu().numpy(), 2.0 * gw[0], rtol=1e-6)
    np.testing.assert_allclose(w1.grad.cpu().numpy(), 2.0 * gw[1], rtol=1e-6)
    np.testing.assert_allclose(s0.grad.cpu().numpy(), 2.0 * gs[0], rtol=1e-6)
    np.testing.assert_allclose(s1.grad.cpu().numpy(), 2.0 * gs[1], rtol=1e-6)
    g1326 = to_combos(spot, 0, gw[0])
    assert g1326.shape == (1326,) and np.count_nonzero(g1326) <= spot.num_hands[0]


def test_device_resident_path_matches_host_path(built):
    """gto_hip_game_value_device / gto_hip_upload_strategy_device: the same
    value and gradients bit for bit as the host-pointer entries, left on the
    GPU; host memory where device memory is required is GTO_ERR_INVALID."""
    import ctypes

    import torch

    from gtopen_amd import _lib as L
    from gtopen_amd import workloads

    _say("case: river spot, device-resident game value vs host path")
    spot, sigma, gv, _ = _solved(workloads.river_spot(), 40)
    for p in range(2):
        u_h, gw_h, gs_h = gv.value_and_grad(p)
        u_d, gw_d, gs_d = gv.value_and_grad_device(p)
        assert u_d == u_h
        for a, b in zip(gw_h + gs_h, gw_d + gs_d):
            assert b.is_cuda and np.array_equal(a, b.cpu().numpy())
    # the strategy from device tensors: a changed sigma moves the value the
    # same way through both uploads
    s0 = sigma[0] * np.float32(0.5)
    gv.set_strategy(s0, sigma[1])
    u_h = gv.value(0)
    gv.set_strategy(*sigma)
    gv.set_strategy_device(torch.tensor(s0, device="cuda"), torch.tensor(sigma[1], device="cuda"))
    u_d, _, _ = gv.value_and_grad_device(0, grad_weights=False, grad_strategy=False)
    assert u_d == u_h
    host = np.zeros(spot.num_hands[0], dtype=np.float32)
    val = ctypes.c_double()
    rc = L.lib().gto_hip_game_value_device(gv.engine._h, 0, ctypes.byref(val), ctypes.c_void_p(host.ctypes.data),
                                           None, None, None)
    assert rc == -1  # GTO_ERR_INVALID
    rc = L.lib().gto_hip_upload_strategy_device(gv.engine._h, ctypes.c_void_p(s0.ctypes.data), None)
    assert rc == -1
    # the engine is still usable after the rejected calls
    assert gv.value(0) == u_h
