"""Differentiable game value over a solved spot.

The forward pass is the average-strategy traversal of the reference
(crates/solver/src/cfr.rs traverse_avg via best_response.rs:128-223) with the
behavioural strategy given as an input, and the value
U_p = sum_i w_p[i] * cfv_p(root)[i] (best_response.rs:264-269). U_p is
linear in each range and in each node's strategy, so its gradient is a set of
counterfactual values; the engine's backward (a second value sweep under the
player's payoffs with the opponent as traverser, then a per-node product of
reach and child values) returns

* dU/dw_p = cfv_p(root), dU/dw_o = root values of o's hands under p's payoffs,
* dU/dsigma_q(node, a, i) = reach_q(node, i) * chance weight * value of child a.

``GameValueFunction`` wraps this in ``torch.autograd.Function``; weights and
strategies travel through the C ABI (gto_hip_set_weights,
gto_hip_upload_arenas, gto_hip_game_value) as host arrays.
"""
import ctypes

import numpy as np

from . import _lib as L
from .solver import GpuSolver


def combo_indices(spot, p):
    """combo_index (cards.rs:100-103) of each of player p's hands."""
    c1 = spot.hand_c1[p].astype(np.int64)
    c2 = spot.hand_c2[p].astype(np.int64)
    hi, lo = np.maximum(c1, c2), np.minimum(c1, c2)
    return hi * (hi - 1) // 2 + lo


def to_combos(spot, p, per_hand):
    """A per-hand vector of player p as a 1326-combo vector (0 off-range)."""
    out = np.zeros(1326, dtype=np.asarray(per_hand).dtype)
    out[combo_indices(spot, p)] = per_hand
    return out


def from_combos(spot, p, per_combo):
    """Player p's hands' entries of a 1326-combo vector."""
    return np.asarray(per_combo)[combo_indices(spot, p)]


def behavioural_strategy(spot, strat):
    """Average strategy of every action node as arena-layout arrays: the
    strategy sums normalised per hand, max(S, 0) / sum over actions with an
    f64 sum, uniform where the sum is <= 1e-12 (cfr.rs:679-702)."""
    out = [np.zeros(spot.data_size[q], dtype=np.float32) for q in range(2)]
    act = np.nonzero(spot.kind == L.KIND_ACTION)[0]
    for n in act:
        q = int(spot.player[n])
        na, nh, off = int(spot.num_children[n]), spot.num_hands[q], int(spot.data_offset[n])
        blk = np.maximum(np.asarray(strat[q][off:off + na * nh], dtype=np.float32), 0.0).reshape(na, nh)
        tot = blk.astype(np.float64).sum(axis=0)
        sig = np.where(tot > 1e-12, blk.astype(np.float64) / np.where(tot > 1e-12, tot, 1.0), 1.0 / na)
        out[q][off:off + na * nh] = sig.astype(np.float32).reshape(-1)
    return out


class GameValue:
    """Game value U_player(w0, w1, sigma0, sigma1) and its gradient on the GPU.

    Holds its own f32 engine for the spot; strategies are arena-layout f32
    arrays (tree data_offset order, data_size[q] entries), weights are
    per-hand arrays in the spot's hand order.
    """

    def __init__(self, spot, device=0):
        self.spot = spot
        # every node's strategy is an input: no suit-orbit sharing
        self.engine = GpuSolver(spot, storage="f32", use_graphs=False, device=device, isomorphism=False)

    def set_weights(self, w0=None, w1=None):
        keep = [None if w is None else np.ascontiguousarray(w, dtype=np.float32) for w in (w0, w1)]
        for q, w in enumerate(keep):
            if w is not None and w.shape != (self.spot.num_hands[q],):
                raise ValueError(f"weights of player {q} need shape ({self.spot.num_hands[q]},), got {w.shape}")
        ptr = [None if w is None else w.ctypes.data_as(L.c_f32p) for w in keep]
        L.check(L.lib().gto_hip_set_weights(self.engine._h, ptr[0], ptr[1]), "set weights")

    def set_strategy(self, sigma0=None, sigma1=None):
        keep = [None if s is None else np.ascontiguousarray(s, dtype=np.float32) for s in (sigma0, sigma1)]
        for q, s in enumerate(keep):
            if s is not None and s.shape != (self.spot.data_size[q],):
                raise ValueError(f"strategy of player {q} needs shape ({self.spot.data_size[q]},), got {s.shape}")
        ptr = [None if s is None else s.ctypes.data_as(L.c_f32p) for s in keep]
        L.check(L.lib().gto_hip_upload_arenas(self.engine._h, None, None, ptr[0], ptr[1]), "upload strategy")

    def value_and_grad(self, player, grad_weights=True, grad_strategy=True, noalloc=False):
        """(U, [dU/dw0, dU/dw1], [dU/dsigma0, dU/dsigma1]); skipped parts are None.
        noalloc: through gto_hip_game_value_noalloc (no allocation in the
        call; strategy arenas of at most 16M entries per player)."""
        sp = self.spot
        gw = [np.zeros(sp.num_hands[q], dtype=np.float32) if grad_weights else None for q in range(2)]
        gs = [np.zeros(sp.data_size[q], dtype=np.float32) if grad_strategy else None for q in range(2)]
        val = ctypes.c_double()

        def p(a):
            return None if a is None else a.ctypes.data_as(L.c_f32p)

        L.check(
            (L.lib().gto_hip_game_value_noalloc if noalloc else L.lib().gto_hip_game_value)(
                self.engine._h, player, ctypes.byref(val), p(gw[0]), p(gw[1]), p(gs[0]), p(gs[1])),
            "game value",
        )
        return val.value, gw, gs

    def set_strategy_device(self, sigma0=None, sigma1=None):
        """set_strategy from float32 torch tensors on the engine's GPU
        (contiguous, data_size[q] entries), copied on the device."""
        ptr = []
        for q, t in enumerate((sigma0, sigma1)):
            if t is not None:
                _check_device_tensor(t, (self.spot.data_size[q],), f"strategy of player {q}")
            ptr.append(None if t is None else ctypes.c_void_p(t.data_ptr()))
        _sync_producer()
        L.check(L.lib().gto_hip_upload_strategy_device(self.engine._h, ptr[0], ptr[1]), "upload strategy (device)")

    def value_and_grad_device(self, player, grad_weights=True, grad_strategy=True):
        """value_and_grad with the gradients left on the GPU: (U, [dU/dw0,
        dU/dw1], [dU/dsigma0, dU/dsigma1]) as float32 torch tensors on the
        engine's device (gto_hip_game_value_device); skipped parts are None."""
        import torch

        sp = self.spot
        dev = torch.device("cuda", self.engine.device)
        gw = [torch.empty(sp.num_hands[q], dtype=torch.float32, device=dev) if grad_weights else None
              for q in range(2)]
        gs = [torch.empty(sp.data_size[q], dtype=torch.float32, device=dev) if grad_strategy else None
              for q in range(2)]
        val = ctypes.c_double()

        def p(t):
            return None if t is None else ctypes.c_void_p(t.data_ptr())

        _sync_producer()
        L.check(L.lib().gto_hip_game_value_device(self.engine._h, player, ctypes.byref(val), p(gw[0]), p(gw[1]),
                                                  p(gs[0]), p(gs[1])), "game value (device)")
        return val.value, gw, gs

    def value(self, player):
        return self.value_and_grad(player, grad_weights=False, grad_strategy=False)[0]


def _check_device_tensor(t, shape, what):
    import torch

    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float32 and t.is_contiguous()):
        raise ValueError(f"{what}: needs a contiguous float32 tensor on the GPU")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: needs shape {tuple(shape)}, got {tuple(t.shape)}")


def _sync_producer():
    """The engine runs on its own stream: work queued on torch's streams (the
    tensors' producers) completes first."""
    import torch

    torch.cuda.synchronize()


def _autograd():
    import torch

    class GameValueFunction(torch.autograd.Function):
        """U_player(w0, w1, sigma0, sigma1) with a hand-written backward.

        forward(ctx, gv, player, w0, w1, sigma0, sigma1): per-hand weights and
        arena-layout strategies (float32 tensors, any device) -> float64
        scalar. The backward is computed by the engine in the same call when
        any input needs a gradient.
        """

        @staticmethod
        def forward(ctx, gv, player, w0, w1, sigma0, sigma1):
            dev = w0.device
            # ranges are small and go through the host; strategies and
            # gradients of GPU tensors stay on the GPU
            gv.set_weights(w0.detach().float().cpu().numpy(), w1.detach().float().cpu().numpy())
            on_gpu = sigma0.is_cuda and sigma1.is_cuda and sigma0.device == sigma1.device
            if on_gpu:
                gv.set_strategy_device(sigma0.detach().float().contiguous(), sigma1.detach().float().contiguous())
            else:
                gv.set_strategy(sigma0.detach().float().cpu().numpy(), sigma1.detach().float().cpu().numpy())
            need_w = ctx.needs_input_grad[2] or ctx.needs_input_grad[3]
            need_s = ctx.needs_input_grad[4] or ctx.needs_input_grad[5]
            if on_gpu:
                u, gw, gs = gv.value_and_grad_device(player, grad_weights=need_w, grad_strategy=need_s)
            else:
                u, gw, gs = gv.value_and_grad(player, grad_weights=need_w, grad_strategy=need_s)
                gw = [None if g is None else torch.from_numpy(g) for g in gw]
                gs = [None if g is None else torch.from_numpy(g) for g in gs]
            grads = []
            for g, like in ((gw[0], w0), (gw[1], w1), (gs[0], sigma0), (gs[1], sigma1)):
                grads.append(None if g is None else g.to(device=like.device, dtype=like.dtype))
            ctx.grads = grads
            return torch.tensor(u, dtype=torch.float64, device=dev)

        @staticmethod
        def backward(ctx, gout):
            g = [None if x is None else x * gout.to(x.dtype) for x in ctx.grads]
            return None, None, g[0], g[1], g[2], g[3]

    return GameValueFunction


def game_value(gv, player, w0, w1, sigma0, sigma1):
    """Differentiable U_player (see GameValueFunction)."""
    return _autograd().apply(gv, player, w0, w1, sigma0, sigma1)
