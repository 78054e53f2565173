"""GPU parity of Monte Carlo resampling (pp_resample) and mcpropagate.

With dyadic coefficients every partial sum is exact, so the intervals do not
depend on the order in which weights are added (test/test_montecarlo.jl:580-
624 uses the same device): the device then equals the oracle bit for bit for
every method, comb offset and draw set. With general coefficients the device
adds weights chunk by chunk and the oracle left to right; the intervals then
agree to the last bits and the survivors are compared up to the rare term
whose interval end falls within rounding of a tooth. mcpropagate equals
propagate below max_size and bounds the sum above it
(test/test_montecarlo.jl:241-276)."""
import math

import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import OracleSum
import pauliprop_hip as pp
from pauliprop_hip.paulis import int_to_limbs

pytestmark = pytest.mark.gpu

SEMI, SYSTEMATIC, MULTINOMIAL = 0, 1, 2
FUNCS = {SEMI: pp.semideterministic_systematic_resample, SYSTEMATIC: pp.systematic_resample,
         MULTINOMIAL: pp.multinomial_resample}


def _pair(nq, coeffs, rng):
    W = pp.key_words(nq)
    terms = set()
    while len(terms) < len(coeffs):
        terms.add(int(rng.integers(1, 1 << min(2 * nq, 62))))
    keys = np.array([int_to_limbs(t, W) for t in sorted(terms)], dtype=np.uint64).reshape(-1, W)
    c = np.asarray(coeffs, dtype=np.float64).reshape(-1, 1)
    dev = pp.HIPPauliSum(nq, capacity=2 * len(coeffs))
    dev.upload(keys, c)
    return dev, OracleSum(W, 1).set(keys, c)


@pytest.mark.parametrize("nq", [30, 60, 120])
@pytest.mark.parametrize("method,squared,calibrate",
                         [(SEMI, False, False), (SYSTEMATIC, False, True), (SYSTEMATIC, False, False),
                          (SYSTEMATIC, True, True), (MULTINOMIAL, False, False), (MULTINOMIAL, True, False)])
def test_dyadic_sums_match_oracle_bitwise(nq, method, squared, calibrate):
    print(f"resample nq={nq} method={method} squared={squared} calibrate={calibrate}", flush=True)
    rng = np.random.default_rng(nq * 10 + method * 3 + squared)
    n = 150000
    coeffs = rng.choice([-1, 1], size=n) * rng.integers(1, 65, size=n) / 8.0
    for target in (n // 10, n // 2, n - 37):
        dev, ref = _pair(nq, coeffs, rng)
        offset = float(rng.random())
        draws = np.sort(rng.random(target)) if method == MULTINOMIAL else None
        sp_dev = dev.resample(target, resample_func=FUNCS[method], squared=squared, calibrate=calibrate,
                              offset=offset, draws=draws)
        sp_ref = ref.resample(method, target, squared=squared, offset=offset, draws=draws, calibrate=calibrate)
        assert sp_dev == sp_ref
        assert_bitwise(*dev.download(), *ref.get(), f"resample method={method} target={target}")


@pytest.mark.parametrize("method", [SEMI, SYSTEMATIC])
def test_general_coefficients_match_oracle(method):
    print(f"resample general coefficients method={method}", flush=True)
    rng = np.random.default_rng(77 + method)
    n = 200000
    coeffs = rng.normal(size=n) * np.exp(1.5 * rng.normal(size=n))
    dev, ref = _pair(80, coeffs, rng)
    offset = float(rng.random())
    sp_dev = dev.resample(n // 4, resample_func=FUNCS[method], offset=offset)
    sp_ref = ref.resample(method, n // 4, offset=offset)
    assert sp_dev == pytest.approx(sp_ref, rel=1e-12)
    kd, cd = dev.download()
    kr, cr = ref.get()
    assert abs(len(kd) - len(kr)) <= 2
    same = {tuple(k): c for k, c in zip(map(tuple, kr), cr[:, 0])}
    matched = sum(1 for k, c in zip(map(tuple, kd), cd[:, 0]) if k in same and abs(same[k] - c) <= 1e-12 * abs(c))
    assert matched >= len(kr) - 2


def test_resample_edge_cases():
    print("resample edge cases", flush=True)
    e = pp.HIPPauliSum(20)
    with pytest.raises(ValueError):
        e.resample(5, offset=0.5)  # as the reference: only the multinomial method grows a sum
    assert e.resample(5, resample_func=pp.multinomial_resample, draws=np.linspace(0.0, 0.9, 5)) == 0.0
    assert len(e) == 0
    s = pp.HIPPauliSum.from_paulisum(pp.PauliSum(20, {1: 1.0, 4: -2.0, 7: 0.5}))
    with pytest.raises(ValueError):
        s.resample(4, resample_func=pp.systematic_resample, offset=0.5)
    with pytest.raises(ValueError):
        s.resample(2, resample_func=pp.semideterministic_systematic_resample, squared=True, offset=0.5)
    with pytest.raises(ValueError):
        s.resample(2, offset=1.5)
    with pytest.raises(ValueError):
        s.resample(2, resample_func=pp.multinomial_resample, draws=[0.7, 0.2])
    b = pp.HIPPauliSum(20, batch=2)
    with pytest.raises(pp._lib.PPError):
        b.resample(1, offset=0.5)
    # the multinomial method may grow a sum (draws with replacement)
    s.resample(6, resample_func=pp.multinomial_resample, draws=np.linspace(0.0, 0.99, 6))
    assert 1 <= len(s) <= 3


def test_mcpropagate_matches_propagate_below_max_size_and_bounds_it_above():
    print("mcpropagate", flush=True)
    nq = 16
    circ = pp.efficientsu2circuit(nq, 3)
    th = np.random.default_rng(2).normal(size=pp.countparameters(circ))
    obs = pp.PauliString(nq, "Z", 5)
    exact = pp.propagate(circ, obs, th, min_abs_coeff=0)
    mc = pp.mcpropagate(circ, obs, th, max_size=10 ** 9, min_abs_coeff=0, rng=np.random.default_rng(0))
    assert_bitwise(*mc.download(), *exact.download(), "mcpropagate below max_size")
    for f in (None, pp.systematic_resample, pp.multinomial_resample):
        r = pp.mcpropagate(circ, obs, th, max_size=200, min_abs_coeff=1e-8, resample_func=f,
                           rng=np.random.default_rng(1))
        assert 0 < len(r) <= 200
    # unbiased: the mean estimate over independent runs approaches the exact overlap
    ex = exact.overlapwithzero()
    est = [pp.mcpropagate(circ, obs, th, max_size=400, min_abs_coeff=0, rng=np.random.default_rng(s)).overlapwithzero()
           for s in range(24)]
    assert abs(np.mean(est) - ex) < 4 * np.std(est) / math.sqrt(len(est)) + 1e-9
