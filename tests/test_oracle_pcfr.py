"""The oracle's PCFR+ (cfr.rs:250-289, 337-348, 386-406), pinned without
retyping its formula:

* the reference's own test (convergence.rs:499-521): PCFR+ brings the small
  flop below 1% pot in 200 iterations, with f32 and with compressed arenas;
* an identity of the algorithm: from zero regrets and predictions an
  update stores prediction = regret = r, so the next matching sees
  max(r,0) + r = 2 max(r,0), which normalises (exactly, in binary) to
  CFR+'s max(r,0). Iteration 1 (both sweeps) and sweep 0 of iteration 2 are
  therefore bit-identical to CFR+. Sweep 1 of iteration 2 matches player 0
  on max(R,0) plus the prediction written in sweep 0, which differs. So
  after two iterations player 0's arenas still equal CFR+'s and player 1's
  do not. A prediction that is ignored, has the wrong sign, or is taken
  from the wrong sweep fails one of these checks.
"""
import numpy as np
import pytest

from test_host_core import _cfg


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _flat(arenas):
    (r0, r1), (s0, s1) = arenas
    return [r0, r1, s0, s1]


@pytest.mark.parametrize("storage", ["f32", "compressed"])
def test_pcfr_plus_converges_on_the_small_flop(built, storage):
    import oracle
    from gtopen_amd import Spot

    spot = Spot(_cfg())
    o = oracle.Oracle(spot, storage=storage, algo="pcfr+")
    o.iterate(200)
    e = o.exploitability() / 60.0 * 100.0
    assert e < 1.0, f"PCFR+ ({storage}) should converge below 1% pot, got {e}%"


def test_pcfr_plus_departs_from_cfr_plus_in_the_second_iterations_second_sweep(built):
    import oracle
    from gtopen_amd import Spot

    spot = Spot(_cfg())
    a = oracle.Oracle(spot, algo="pcfr+")
    b = oracle.Oracle(spot, algo="cfr+")
    a.iterate(1)
    b.iterate(1)
    assert _same(_flat(a.arenas()), _flat(b.arenas()))
    a.iterate(1)
    b.iterate(1)
    ra, rb = _flat(a.arenas()), _flat(b.arenas())
    # player 0's regrets and strategy sums (updated in sweep 0) still agree
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[2], rb[2])
    # player 1's (updated in sweep 1, against the predicted strategy) do not
    assert not (np.array_equal(ra[1], rb[1]) and np.array_equal(ra[3], rb[3]))


def test_algorithm_names_parse_like_the_reference(built):
    from gtopen_amd import _lib as L

    for name, code in (("dcfr", 0), ("DCFR", 0), ("cfr+", 1), ("cfr_plus", 1), ("CFR-Plus", 1), ("cfr", 1),
                       ("pcfr+", 2), ("PCFR", 2), ("p-cfr+", 2)):
        assert L.algo_code(name) == code
    with pytest.raises(ValueError):
        L.algo_code("mccfr")


def test_compressed_pcfr_plus_tracks_cfr_plus_then_departs(built):
    """Compressed CFR+ matches on integer sums of the i16 quants
    (cfr.rs:899-919) while compressed PCFR+ decodes to f32 first
    (cfr.rs:276-288), so the doubling identity holds only to f32 rounding:
    after one iteration the two agree closely; after three, the prediction
    has moved PCFR+ away."""
    import oracle
    from gtopen_amd import Spot

    spot = Spot(_cfg())
    a = oracle.Oracle(spot, storage="compressed", algo="pcfr+")
    b = oracle.Oracle(spot, storage="compressed", algo="cfr+")
    a.iterate(1)
    b.iterate(1)
    for x, y in zip(_flat(a.arenas()), _flat(b.arenas())):
        assert np.allclose(x, y, rtol=1e-3, atol=1e-3 * max(1.0, float(np.abs(y).max())))
    a.iterate(2)
    b.iterate(2)
    diff = max(float(np.abs(x - y).max()) for x, y in zip(_flat(a.arenas()), _flat(b.arenas())))
    scale = max(float(np.abs(y).max()) for y in _flat(b.arenas()))
    assert diff > 1e-2 * scale
