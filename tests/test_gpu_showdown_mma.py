"""The matrix-core showdown (BF16 and FP8 precision forms) against the exact
showdown. With weights and strategies whose reach values are exactly
representable in the operand format the forms must agree with the exact
kernel to f32 accumulation rounding; with general inputs, to the operand
rounding of the reach (bf16: 8 significant bits, e4m3: 4). The values are
those of both players' current strategies (the values a DCFR update sees);
exploitability always uses the exact showdown."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CURRENT = 2  # gto_hip_node_values mode: current strategies


def _say(msg):
    print(msg, flush=True)


def _pow2_regrets(spot):
    """Regrets whose regret-matched strategy is (1, 1, 2, 4, ...) / 2^(na-1)
    at every node, the same for every hand: reach rows are the range
    weights times a power of two."""
    from gtopen_amd import _lib as L

    r = [np.zeros(spot.data_size[q], np.float32) for q in range(2)]
    for n in np.nonzero(spot.kind == L.KIND_ACTION)[0]:
        q, na = int(spot.player[n]), int(spot.num_children[n])
        nh, off = spot.num_hands[q], int(spot.data_offset[n])
        for a in range(na):
            r[q][off + a * nh:off + (a + 1) * nh] = float(2 ** max(0, a - 1))
    return r


def _bf16_grid(w):
    return (np.asarray(w, np.float32).view(np.uint32) & np.uint32(0xFFFF0000)).view(np.float32)


def _e4m3_grid(w):
    """Weights clipped to [0.25, 1] and rounded to 4 significant bits: after
    the per-node power-of-two scale every reach value is a normal e4m3."""
    m, e = np.frexp(np.clip(np.asarray(w, np.float64), 0.25, 1.0))
    return np.ldexp(np.round(m * 16.0) / 16.0, e).astype(np.float32)


def _rows(spot, mode, weights, regrets, show):
    from gtopen_amd import GpuSolver, _lib as L

    g = GpuSolver(spot, showdown=mode)
    g.upload_arenas(regrets, None)
    L.check(L.lib().gto_hip_set_weights(g._h, weights[0].ctypes.data_as(L.c_f32p),
                                        weights[1].ctypes.data_as(L.c_f32p)), "set weights")
    rows = [g.node_values(p, CURRENT, show) for p in range(2)]
    g.close()
    return rows


@pytest.mark.parametrize("form,grid", [("bf16", _bf16_grid), ("fp8", _e4m3_grid)])
def test_mma_showdown_exact_on_the_operand_grid(built, form, grid):
    from gtopen_amd import Spot, _lib as L, workloads

    for name, cfg in (("river", workloads.river_spot()), ("turn", workloads.turn_spot())):
        _say(f"case: {name} spot, reach on the {form} grid: MFMA showdown rows vs exact rows")
        spot = Spot(cfg)
        show = np.nonzero(spot.kind == L.KIND_SHOWDOWN)[0]
        regrets = _pow2_regrets(spot)
        w = [grid(spot.weights[q]) for q in range(2)]
        ex = _rows(spot, "exact", w, regrets, show)
        mm = _rows(spot, form, w, regrets, show)
        for p in range(2):
            a, b = ex[p].astype(np.float64), mm[p].astype(np.float64)
            scale = np.abs(a).max()
            err = np.abs(a - b).max()
            _say(f"  traverser {p}: {len(show)} showdown rows, max |mma - exact| {err:.3e} (scale {scale:.3e})")
            assert scale > 0
            assert err <= 1e-5 * scale


@pytest.mark.parametrize("form,bits", [("bf16", 8), ("fp8", 4)])
def test_mma_showdown_general_inputs(built, form, bits):
    from gtopen_amd import GpuSolver, Spot, _lib as L, workloads

    _say(f"case: river and turn spots after 30 f32 iterations: {form} MFMA vs exact showdown rows")
    for cfg in (workloads.river_spot(), workloads.turn_spot()):
        spot = Spot(cfg)
        show = np.nonzero(spot.kind == L.KIND_SHOWDOWN)[0]
        g = GpuSolver(spot)
        g.iterate(30)
        ex = [g.node_values(p, CURRENT, show) for p in range(2)]
        g.set_showdown(form)
        mm = [g.node_values(p, CURRENT, show) for p in range(2)]
        # the average-strategy values stay exact under a precision form
        avg_mm = g.node_values(0, 1, show)
        g.set_showdown("exact")
        avg_ex = g.node_values(0, 1, show)
        g.close()
        assert np.array_equal(avg_mm, avg_ex)
        for p in range(2):
            a, b = ex[p].astype(np.float64), mm[p].astype(np.float64)
            scale = np.abs(a).max()
            err = np.abs(a - b).max()
            _say(f"  {spot.num_nodes} nodes, traverser {p}: max |mma - exact| {err:.3e} (scale {scale:.3e})")
            # round-to-nearest keeps each reach value within 2^-bits relative
            assert err <= 2.0 ** -bits * scale


@pytest.mark.parametrize("form,tol", [("bf16", 1e-3), ("fp8", 1e-2)])
def test_mma_showdown_trajectory(built, form, tol):
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads

    _say(f"case: turn spot, {form} showdown engine vs FP64 oracle, 150 iterations")
    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot, showdown=form)
    ref = oracle.Oracle(spot, precision=64)
    pot = spot.starting_pot
    for k in range(6):
        g.iterate(25)
        ref.iterate(25)
        eg, e64 = g.exploitability(), ref.exploitability()
        _say(f"  iter {25 * (k + 1)}: {form} {eg:.6f} fp64 {e64:.6f} ({abs(eg - e64) / pot * 100:.4f}% pot)")
        assert abs(eg - e64) <= tol * pot
    g.close()
