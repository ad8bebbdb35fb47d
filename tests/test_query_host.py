"""Host-side lock construction (crates/solver/src/query.rs:40-150): the
frequency normaliser and the proportional-fitting rake. Checked by the
properties that define them, not by re-deriving their formulas."""
import numpy as np
import pytest


def test_normalize_freqs(built):
    from gtopen_amd.query import QueryError, normalize_freqs

    np.testing.assert_allclose(normalize_freqs([2.0, 6.0], 2), [0.25, 0.75], rtol=1e-7)
    for bad, msg in (([1.0], "expected 2"), ([-1.0, 2.0], "finite"), ([np.inf, 1.0], "finite"),
                     ([0.0, 0.0], "positive"), ([3e38, 3e38], "not a frequency vector")):
        with pytest.raises(QueryError, match=msg):
            normalize_freqs(bad, 2)


def _random_strategy(rng, na, nh):
    s = rng.random((na, nh)).astype(np.float32) ** 3
    return (s / s.sum(axis=0)).astype(np.float32)


def _aggregate(sig, reach):
    w = reach.astype(np.float64)
    return (sig.astype(np.float64) * w).sum(axis=1) / w.sum()


def test_rake_hits_target_with_one_multiplier_per_action(built):
    from gtopen_amd.query import rake_to_target

    rng = np.random.default_rng(1)
    for na in (2, 3, 4):
        sig = _random_strategy(rng, na, 300)
        reach = rng.random(300).astype(np.float32)
        target = rng.random(na).astype(np.float32)
        target /= target.sum()
        out = rake_to_target(sig, reach, target)
        np.testing.assert_allclose(out.sum(axis=0), 1.0, atol=1e-5)
        assert np.max(np.abs(_aggregate(out, reach) - target)) < 2e-4
        # out = diag(m) sig, renormalised per hand: the ratio out/sig of two
        # actions is the same for every hand
        r = (out[0].astype(np.float64) / sig[0]) / (out[1].astype(np.float64) / sig[1])
        assert np.ptp(r) / np.mean(r) < 1e-5
        # same base, same target: the same result (no compounding)
        assert np.array_equal(rake_to_target(sig, reach, target), out)


def test_rake_zero_target_and_massless_hands(built):
    from gtopen_amd.query import rake_to_target

    rng = np.random.default_rng(2)
    sig = _random_strategy(rng, 3, 50)
    sig[1:, :5] = 0.0  # five hands play only action 0 ...
    sig[0, :5] = 1.0
    reach = np.ones(50, np.float32)
    target = np.array([0.0, 0.4, 0.6], np.float32)  # ... which the target forbids
    out = rake_to_target(sig, reach, target)
    assert np.all(out[0] == 0.0)
    # massless hands start from the target itself and are then scaled like
    # every other hand: identical columns, nothing on the forbidden action
    assert np.all(out[:, :5] == out[:, :1]) and out[1, 0] > 0 and out[2, 0] > 0
    assert np.max(np.abs(_aggregate(out, reach) - target)) < 2e-4
