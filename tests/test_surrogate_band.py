"""The surrogate's band calibration (qubatron_amd.surrogate.choose_band), on
the CPU: the band it returns keeps the decided-and-wrong queries within the
target, no band of its grid decides more within the target (brute force over
the same grid), and the decision rule is the render path's strict p < lo /
p > hi."""
import numpy as np
import pytest

from qubatron_amd import surrogate as S


def _score(p, y, lo, hi):
    d0, d1 = p < lo, p > hi
    return int(np.count_nonzero(d0) + np.count_nonzero(d1)), int(np.count_nonzero(d0 & y) + np.count_nonzero(d1 & ~y))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_band_meets_target_and_is_best_on_its_grid(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    y = rng.random(n) < 0.35
    # scores: informative but noisy, with a hard middle
    p = np.clip(np.where(y, rng.beta(4, 1.5, n), rng.beta(1.5, 4, n)), 0.0, 1.0)
    target = 0.01
    lo, hi, dec, mis = S.choose_band(p, y, target, steps=40)
    nd, ne = _score(p, y, lo, hi)
    assert ne <= target * n and abs(mis - ne / n) < 1e-12 and abs(dec - nd / n) < 1e-12
    best = 0
    for a in np.linspace(0, 0.5, 41)[1:]:
        for b in np.linspace(0.5, 1, 41)[:-1]:
            d, e = _score(p, y, a, b)
            if e <= target * n:
                best = max(best, d)
    assert nd == best


def test_separable_scores_decide_everything():
    p = np.array([0.01, 0.02, 0.2, 0.8, 0.97, 0.99])
    y = np.array([0, 0, 0, 1, 1, 1], bool)
    lo, hi, dec, mis = S.choose_band(p, y, 0.0, steps=100)
    assert dec == 1.0 and mis == 0.0 and 0.2 < lo <= 0.5 <= hi < 0.8


def test_no_band_within_target_decides_almost_nothing():
    p = np.array([0.0001, 0.9999, 0.5])
    y = np.array([1, 0, 1], bool)  # the confident ones are wrong
    lo, hi, dec, mis = S.choose_band(p, y, 0.0, steps=10)
    assert lo == 1e-6 and hi == 1.0 - 1e-6 and dec == 0.0 and mis == 0.0
