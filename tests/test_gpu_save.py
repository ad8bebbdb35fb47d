"""GTOSOLVE2 round trips from GPU arenas (gpu/mod.rs:845-867 sync_to_cpu,
then save.rs): f32 and compressed engines, a GPU save read by the CPU
oracle, and locks riding along in the header."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def test_f32_gpu_roundtrip(built, tmp_path):
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.save import load_solver

    _say("case: river spot, f32 GPU engine, 50 iterations, save -> load -> compare")
    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot)
    g.iterate(50)
    e0 = g.exploitability()
    path = str(tmp_path / "f32.gto")
    g.save(path)
    spot2, g2, sf = load_solver(path)
    assert sf.iteration == 50 and g2.iteration == 50
    a, b = g.download_arenas(), g2.download_arenas()
    for x, y in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(x, y)
    assert abs(g2.exploitability() - e0) < 1e-9
    # both engines continue identically
    g.iterate(10)
    g2.iterate(10)
    assert g.exploitability() == g2.exploitability()


def test_compressed_gpu_roundtrip(built, tmp_path):
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.save import load_solver

    _say("case: river spot, compressed GPU engine, save -> compressed and f32 loads")
    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot, storage="compressed")
    g.iterate(50)
    e0 = g.exploitability()
    path = str(tmp_path / "c.gto")
    g.save(path)
    _, gc, _ = load_solver(path, storage="compressed")
    ec = gc.exploitability()
    _say(f"  exploitability: before {e0:.9f}, compressed reload {ec:.9f}")
    assert abs(ec - e0) < 1e-9
    for p in range(2):
        q0, q1 = g.download_compressed(p), gc.download_compressed(p)
        assert np.array_equal(q0[1], q1[1]), "strategy quants changed on re-encode"
        assert np.array_equal(q0[0], q1[0]), "regret quants changed on re-encode"
    _, gf, _ = load_solver(path, storage="f32")
    assert abs(gf.exploitability() - e0) < 1e-4


def test_gpu_save_read_by_oracle(built, tmp_path):
    import oracle
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.save import SaveFile

    _say("case: turn spot, GPU save read back by the CPU oracle")
    spot = Spot(workloads.turn_spot())
    g = GpuSolver(spot)
    g.iterate(20)
    path = str(tmp_path / "turn.gto")
    g.save(path)
    sf = SaveFile(path)
    r, s = sf.read_arenas(spot)
    o = oracle.Oracle(spot)
    o.set_arenas(r, s)
    o.iteration = sf.iteration
    eg, ec = g.exploitability(), o.exploitability()
    _say(f"  exploitability gpu {eg:.9f} oracle {ec:.9f}")
    assert abs(eg - ec) <= 1e-6 * spot.starting_pot


def test_locks_ride_along(built, tmp_path):
    from gtopen_amd import GpuSolver, Spot, workloads
    from gtopen_amd.save import load_solver

    _say("case: river spot, root lock saved in the header and re-applied on load")
    spot = Spot(workloads.river_spot())
    g = GpuSolver(spot)
    g.iterate(20)
    na, nh = int(spot.num_children[0]), spot.num_hands[int(spot.player[0])]
    sig = np.zeros((na, nh), np.float32)
    sig[0] = 0.7
    sig[1] = 0.3
    g.update_locks({0: sig.reshape(-1)})
    g.lock_labels = {0: "root 30% bet"}
    g.iterate(10)
    path = str(tmp_path / "locks.gto")
    g.save(path)
    _, g2, sf = load_solver(path)
    assert list(sf.locks) == [0] and np.array_equal(sf.locks[0], sig.reshape(-1))
    assert sf.labels == {0: "root 30% bet"}
    g.iterate(10)
    g2.iterate(10)
    a, b = g.download_arenas(), g2.download_arenas()
    for x, y in zip(a[0] + a[1], b[0] + b[1]):
        assert np.array_equal(x, y)
