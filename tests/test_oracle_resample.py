"""The oracle's Monte Carlo resampling against the reference's own tests of
src/Base/MonteCarlo/resample.jl (test/test_montecarlo.jl:320-662): the mean
over every comb offset is the incoming coefficients, the calibrated comb
misses its expectations by less than one tooth, the calibrated spacing keeps
the target in expectation, heavy terms survive as they are, and multinomial
draws land in the intervals a brute-force count finds."""
import math

import numpy as np
import pytest

from oracle import OracleSum

SEMI, SYSTEMATIC, MULTINOMIAL = 0, 1, 2


def _sum(coeffs):
    n = len(coeffs)
    keys = np.arange(1, n + 1, dtype=np.uint64).reshape(-1, 1)
    return OracleSum(1, 1).set(keys, np.asarray(coeffs, dtype=np.float64).reshape(-1, 1))


def _as_dict(s):
    k, c = s.get()
    return {int(a[0]): float(b[0]) for a, b in zip(k, c)}


@pytest.mark.parametrize("method", [SYSTEMATIC, SEMI])
def test_mean_over_every_comb_offset_is_the_incoming_coefficients(method):
    # test/test_montecarlo.jl:526-554: between two offsets where a tooth crosses an
    # interval end the result does not change, so one run per stretch, weighted
    # by its length, gives the exact mean
    n = 40
    rng = np.random.default_rng(9)
    coeffs = rng.normal(size=n) * np.exp(rng.normal(size=n))
    target = 15
    probe = _sum(coeffs)
    spacing = probe.resample(method, target, offset=0.0, calibrate=False)
    if method == SEMI:
        thr = np.abs(coeffs).sum() / target
        w = np.where(np.abs(coeffs) > thr, 0.0, np.abs(coeffs))
    else:
        w = np.abs(coeffs)
    ends = np.cumsum(w)
    edges = np.unique(np.sort(np.concatenate([[0.0], np.mod(ends / spacing, 1.0), [1.0]])))
    mean = np.zeros(n)
    for lo, hi in zip(edges[:-1], edges[1:]):
        s = _sum(coeffs)
        s.resample(method, target, offset=(lo + hi) / 2, calibrate=False)
        for t, c in _as_dict(s).items():
            mean[t - 1] += (hi - lo) * c
    np.testing.assert_allclose(mean, coeffs, rtol=1e-12)


def test_calibrated_comb_misses_its_expectations_by_less_than_one_tooth():
    # test/test_montecarlo.jl:556-578 (heaviest first)
    n = 1000
    rng = np.random.default_rng(11)
    c = rng.normal(size=n) * np.exp(rng.normal(size=n))
    coeffs = c[np.argsort(-np.abs(c), kind="stable")]
    total = np.abs(coeffs).sum()
    for offset in (0.0, 0.3, 0.77, 0.999):
        s = _sum(coeffs)
        spacing = s.resample(SYSTEMATIC, n // 2, offset=offset)
        expected = np.minimum(1.0, np.abs(coeffs) / spacing).sum()
        got = _as_dict(s)
        assert abs(sum(abs(v) for v in got.values()) - total) <= (1 + 1e-9) * spacing
        assert math.floor(expected) <= len(got) <= math.ceil(expected)


@pytest.mark.parametrize("frac", [10, 2, 100])
def test_calibrated_spacing_keeps_the_target_in_expectation(frac):
    # test/test_montecarlo.jl:626-643: log-normal magnitudes over several decades
    n = 4096
    rng = np.random.default_rng(5)
    coeffs = rng.normal(size=n) * np.exp(1.5 * rng.normal(size=n))
    target = n // frac if frac != 100 else n - n // 100
    s = _sum(coeffs)
    spacing = s.resample(SYSTEMATIC, target, offset=0.5, atol=0.0)
    expected = np.minimum(1.0, np.abs(coeffs) / spacing).sum()
    assert 0.99 * target <= expected <= target * (1 + 1e-9)


def test_semideterministic_keeps_heavy_terms_as_they_are():
    # resample.jl:139-170: terms above sum|c| / target_size survive unchanged
    n = 500
    rng = np.random.default_rng(12)
    coeffs = rng.normal(size=n) * np.exp(1.2 * rng.normal(size=n))
    target = 100
    thr = np.abs(coeffs).sum() / target
    s = _sum(coeffs)
    s.resample(SEMI, target, offset=0.42)
    got = _as_dict(s)
    for i in np.nonzero(np.abs(coeffs) > thr)[0]:
        assert got[i + 1] == coeffs[i]
    assert len(got) <= target + 3


@pytest.mark.parametrize("squared", [False, True])
def test_multinomial_draws_land_in_their_intervals(squared):
    n, target = 300, 120
    rng = np.random.default_rng(13)
    coeffs = rng.normal(size=n)
    draws = np.sort(rng.random(target))
    s = _sum(coeffs)
    w_per = s.resample(MULTINOMIAL, target, squared=squared, draws=draws)
    w = coeffs ** 2 if squared else np.abs(coeffs)
    total = math.fsum(w)
    scaled = draws * np.sum(w)
    ends = np.cumsum(w)
    starts = ends - w
    got = _as_dict(s)
    assert len(got) <= target
    for i in range(n):
        k = int(np.sum((scaled >= starts[i]) & (scaled < ends[i])))
        want = math.sqrt(k * w_per) if squared else math.copysign(k * w_per, coeffs[i])
        assert got.get(i + 1, 0.0) == pytest.approx(want, rel=1e-12, abs=0.0)
    assert w_per == pytest.approx(total / target, rel=1e-12)


def test_resampling_argument_errors():
    s = _sum([1.0, -2.0, 0.5])
    with pytest.raises(ValueError):
        s.resample(SYSTEMATIC, 4)          # cannot grow without multinomial draws
    with pytest.raises(ValueError):
        s.resample(SEMI, 2, squared=True)  # squared is multinomial / systematic only
