"""GPU parity of the readouts (pp_overlap, pp_get_coeff, pp_scalar_product):
against the oracle's restatement on random sums of every key width, batched
and not, on the sum a propagation leaves, on the reference's known answers
(test/test_overlaps.jl) and on edge cases (empty sums, absent strings, bad
arguments). Sums are accumulated in double-double on the device and in
array order in the oracle: agreement within 1e-12 of the magnitude."""
import numpy as np
import pytest

from oracle import OracleSum, oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.paulis import int_to_limbs, symboltoint

pytestmark = pytest.mark.gpu


def _rand_terms(rng, nq, n):
    terms = set()
    while len(terms) < n:
        t = 0
        for q in rng.choice(nq, size=int(rng.integers(0, min(nq, 8) + 1)), replace=False):
            t |= int(rng.integers(1, 4)) << (2 * int(q))
        terms.add(t)
    return sorted(terms)


def _pair(nq, terms, coeffs):
    """The same sum on the device and in the oracle."""
    W = pp.key_words(nq)
    B = coeffs.shape[1]
    keys = np.array([int_to_limbs(t, W) for t in terms], dtype=np.uint64).reshape(-1, W)
    dev = pp.HIPPauliSum(nq, batch=B, capacity=max(1024, 2 * len(terms)))
    dev.upload(keys, coeffs)
    ref = OracleSum(W, B).set(keys, coeffs)
    return dev, ref, W


def _close(got, want):
    got = np.atleast_1d(np.asarray(got, dtype=np.float64))
    want = np.atleast_1d(want)
    scale = max(1.0, float(np.abs(want).max()))
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("nq,B", [(20, 1), (50, 1), (120, 1), (50, 3), (120, 2)])
def test_readouts_match_oracle(nq, B):
    print(f"readouts nq={nq} B={B}", flush=True)
    rng = np.random.default_rng(nq * 7 + B)
    terms = _rand_terms(rng, nq, 30000)
    if 0 not in terms:
        terms = [0] + terms
    coeffs = rng.normal(size=(len(terms), B))
    dev, ref, W = _pair(nq, terms, coeffs)
    _close(dev.overlapwithzero(), ref.overlap_kind(0))
    _close(dev.overlapwithplus(), ref.overlap_kind(1))
    for k in range(4):
        ones = sorted(rng.choice(np.arange(1, nq + 1), size=int(rng.integers(0, nq)), replace=False).tolist())
        m = np.array(int_to_limbs(sum(1 << (2 * (q - 1)) for q in ones), W), dtype=np.uint64)
        _close(dev.overlapwithcomputational(ones), ref.overlap_kind(2, m))
    # present strings, absent strings and the identity
    for t in [terms[0], terms[len(terms) // 2], terms[-1]] + [int(x) for x in rng.choice(terms, 5)]:
        _close(dev.getcoeff(t), ref.get_coeff(np.array(int_to_limbs(t, W), dtype=np.uint64)))
    present = set(terms)
    absent = next(t for t in range(1, 1 << 20) if t not in present and t >> (2 * nq) == 0)
    _close(dev.getcoeff(absent), np.zeros(B))
    _close(dev.overlapwithmaxmixed(), ref.get_coeff(np.zeros(W, dtype=np.uint64)))
    # scalar product with a second sum sharing part of the strings
    terms2 = sorted(set(rng.choice(terms, 12000).tolist()) | set(_rand_terms(rng, nq, 5000)))
    coeffs2 = rng.normal(size=(len(terms2), B))
    dev2, ref2, _ = _pair(nq, terms2, coeffs2)
    _close(dev.scalarproduct(dev2), ref.scalar_product(ref2))
    _close(dev2.scalarproduct(dev), ref2.scalar_product(ref))
    _close(pp.overlapwithpaulisum(dev, dev2), ref.scalar_product(ref2) * 2.0 ** nq)


def test_readouts_of_a_propagated_sum():
    print("readouts after a propagation", flush=True)
    nq = 24
    circ = pp.tfitrottercircuit(nq, 6, topology=pp.bricklayertopology(nq, periodic=True))
    th = np.random.default_rng(5).uniform(0.05, 0.4, size=pp.countparameters(circ))
    obs = pp.PauliSum(nq)
    obs.add(["Z"], [12])
    obs.add(["X", "X"], [3, 4], 0.5)
    dev = pp.propagate(circ, obs, th, max_weight=6, min_abs_coeff=1e-6)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=6, min_abs_coeff=1e-6)
    W = pp.key_words(nq)
    _close(dev.overlapwithplus(), ref.overlap_kind(1))
    ones = [1, 5, 12, 13, 20]
    m = np.array(int_to_limbs(sum(1 << (2 * (q - 1)) for q in ones), W), dtype=np.uint64)
    _close(dev.overlapwithcomputational(ones), ref.overlap_kind(2, m))
    _close(dev.overlapwithmaxmixed(), ref.get_coeff(np.zeros(W, dtype=np.uint64)))
    k, _ = dev.download()
    t = int(k[len(k) // 3, 0])
    _close(dev.getcoeff(t), ref.get_coeff(np.array([t], dtype=np.uint64)))
    _close(dev.scalarproduct(dev), ref.scalar_product(ref))


def test_known_answers_on_device():
    # test/test_overlaps.jl:9-111 through the device
    print("readout known answers", flush=True)
    for nq, idx, want in ((4, (2, 4), 1.0), (5, (2,), -1.0), (4, (3,), 1.0), (4, (), 1.0)):
        s = pp.HIPPauliSum.from_paulisum(pp.PauliString(nq, ["Z", "Z"], [2, 4]))
        assert s.overlapwithcomputational(idx) == want
        s = pp.HIPPauliSum.from_paulisum(pp.PauliString(nq, ["X", "Z", "Z"], [1, 2, 4]))
        assert s.overlapwithcomputational(idx) == 0.0
    p1 = pp.PauliString(4, ["I", "Z"], [1, 2], 1.0)
    p2 = pp.PauliString(4, ["X", "Y"], [2, 3], -2.0)
    p3 = pp.PauliString(4, ["Z", "X"], [3, 4], 0.2)
    ps = pp.PauliSum(4, {p1.term: p1.coeff, p2.term: p2.coeff})
    dev = pp.HIPPauliSum.from_paulisum(ps)
    assert pp.scalarproduct(dev, p3) == 0.0 and pp.scalarproduct(p3, dev) == 0.0
    assert pp.scalarproduct(dev, p1) == 1.0 and pp.scalarproduct(dev, p2) == 4.0
    assert pp.scalarproduct(dev, dev) == 5.0
    nq = 3
    q = pp.PauliSum(nq)
    q.add(["I"], [1], 0.2)
    q.add(["X"], [2], 0.3)
    q.add(["Z"], [3], 0.4)
    q.add(["Z", "Z"], [1, 2], 0.5)
    dq = pp.HIPPauliSum.from_paulisum(q)
    rho = pp.HIPPauliSum.from_paulisum(pp.PauliString(nq, "I", 1, 1 / 2 ** nq))
    assert pp.overlapwithpaulisum(rho, dq) == pytest.approx(0.2, abs=1e-15)
    assert pp.overlapwithmaxmixed(dq) == 0.2
    assert pp.overlapwithplus(rho) == 1 / 2 ** nq
    assert pp.overlapwithzero(dq) == pytest.approx(1.1, abs=1e-15)


def test_readout_edge_cases():
    print("readout edge cases", flush=True)
    e = pp.HIPPauliSum(40, batch=2)
    assert np.array_equal(e.overlapwithplus(), np.zeros(2))
    assert np.array_equal(e.overlapwithcomputational([1, 40]), np.zeros(2))
    assert np.array_equal(e.getcoeff(0), np.zeros(2))
    f = pp.HIPPauliSum(40, batch=2)
    assert np.array_equal(e.scalarproduct(f), np.zeros(2))
    with pytest.raises(ValueError):
        e.overlapwithcomputational([41])
    with pytest.raises(ValueError):
        e.scalarproduct(pp.HIPPauliSum(41, batch=2))
    with pytest.raises(ValueError):
        e.scalarproduct(pp.HIPPauliSum(40, batch=1))
