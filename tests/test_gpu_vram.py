"""estimate_vram (gpu/mod.rs:946) must bound what an engine allocates: the
report runner packs waves of flops by it. Measured as the drop in free
device memory across engine creation, after a first engine has set up the
device context."""
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_estimate_bounds_the_allocation(built):
    from gtopen_amd import GpuSolver, Spot, _lib as L, workloads
    from gtopen_amd.solver import vram_info

    warm = GpuSolver(Spot(workloads.river_spot()))
    warm.iterate(1)
    cfgs = [("river", workloads.river_spot()), ("turn", workloads.turn_spot()),
            ("batched flop", workloads.batched_flops(n=1)[0])]
    for name, cfg in cfgs:
        spot = Spot(cfg)
        for storage in ("f32", "compressed", "half"):
            est = L.lib().gto_hip_estimate_vram(spot.view_ptr, L.STORAGE[storage]) / 2 ** 20
            free0, _ = vram_info(0)
            g = GpuSolver(spot, storage=storage)
            g.iterate(2)
            g.synchronize()
            free1, _ = vram_info(0)
            used = free0 - free1
            _say(f"  {name} ({spot.num_nodes} nodes), {storage}: allocated {used} MiB, estimate {est:.0f} MiB")
            g.close()
            assert used <= est
    warm.close()
