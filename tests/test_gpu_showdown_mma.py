"""The matrix-core showdown (BF16 precision form) against the exact
showdown. With weights and strategies whose reach values are exactly
representable in bf16 the two must agree to f32 accumulation rounding;
with general inputs, to the bf16 rounding of the reach."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _pow2_strategy(spot):
    """Strategy sums whose normalised strategy is (1, 1, 2, 4, ...) / 2^(na-1)
    at every node: powers of two, so reach products stay on the bf16 grid."""
    from gtopen_amd import _lib as L

    s = [np.zeros(spot.data_size[q], np.float32) for q in range(2)]
    for n in np.nonzero(spot.kind == L.KIND_ACTION)[0]:
        q, na = int(spot.player[n]), int(spot.num_children[n])
        nh, off = spot.num_hands[q], int(spot.data_offset[n])
        for a in range(na):
            s[q][off + a * nh:off + (a + 1) * nh] = float(2 ** max(0, a - 1))
    return s


def _bf16_grid(w):
    return (np.asarray(w, np.float32).view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def test_mma_showdown_exact_on_the_bf16_grid(built):
    from gtopen_amd import GpuSolver, Spot, _lib as L, workloads

    _say("case: turn spot, reach on the bf16 grid: MFMA showdown rows vs exact rows")
    spot = Spot(workloads.turn_spot())
    show = np.nonzero(spot.kind == L.KIND_SHOWDOWN)[0]
    strat = _pow2_strategy(spot)
    w = [_bf16_grid(spot.weights[q]) for q in range(2)]
    rows = {}
    for mode in ("exact", "bf16"):
        g = GpuSolver(spot, showdown=mode)
        g.upload_arenas(None, strat)
        L.check(L.lib().gto_hip_set_weights(g._h, w[0].ctypes.data_as(L.c_f32p), w[1].ctypes.data_as(L.c_f32p)),
                "set weights")
        rows[mode] = [g.node_values(p, 1, show) for p in range(2)]
    for p in range(2):
        a, b = rows["exact"][p].astype(np.float64), rows["bf16"][p].astype(np.float64)
        scale = np.abs(a).max()
        err = np.abs(a - b).max()
        _say(f"  traverser {p}: {len(show)} showdown rows, max |mma - exact| {err:.3e} (scale {scale:.3e})")
        assert err <= 1e-5 * scale


def test_mma_showdown_general_inputs(built):
    from gtopen_amd import GpuSolver, Spot, _lib as L, workloads

    _say("case: river and turn spots after 30 f32 iterations: MFMA vs exact showdown rows")
    for cfg in (workloads.river_spot(), workloads.turn_spot()):
        spot = Spot(cfg)
        show = np.nonzero(spot.kind == L.KIND_SHOWDOWN)[0]
        g = GpuSolver(spot)
        g.iterate(30)
        r, s = g.download_arenas()
        ex = [g.node_values(p, 1, show) for p in range(2)]
        g.set_showdown("bf16")
        mm = [g.node_values(p, 1, show) for p in range(2)]
        for p in range(2):
            a, b = ex[p].astype(np.float64), mm[p].astype(np.float64)
            scale = np.abs(a).max()
            err = np.abs(a - b).max()
            _say(f"  {spot.num_nodes} nodes, traverser {p}: max |mma - exact| {err:.3e} (scale {scale:.3e})")
            # bf16 keeps 8 significant bits of each reach value: 2^-8 relative
            assert err <= 2.0 ** -8 * scale


def test_mma_showdown_trajectory(built):
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say("case: turn spot, BF16 showdown engine vs FP64 oracle, 150 iterations")
    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot, showdown="bf16")
    ref = oracle.Oracle(spot, precision=64)
    pot = spot.starting_pot
    for k in range(6):
        g.iterate(25)
        ref.iterate(25)
        eg, e64 = g.exploitability(), ref.exploitability()
        _say(f"  iter {25 * (k + 1)}: bf16 {eg:.6f} fp64 {e64:.6f} ({abs(eg - e64) / pot * 100:.4f}% pot)")
        assert abs(eg - e64) <= 1e-3 * pot
