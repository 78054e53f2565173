"""The oracle's readouts (overlap with |0..0>, |+..+>, computational basis
states and the maximally mixed state, getcoeff, scalarproduct) against the
reference's known answers (test/test_overlaps.jl) and against dense matrices:
<psi|O|psi> and Tr[O]/2^n, Tr[A B]/2^n computed from 2^n x 2^n operators."""
import numpy as np
import pytest

from oracle import OracleSum
from dense import operator_from_terms
from pauliprop_hip.paulis import int_to_limbs, symboltoint


def _sum(nq, terms):
    """terms: {int term: coeff} -> OracleSum (W for nq qubits)."""
    W = 1 if nq <= 32 else (2 if nq <= 64 else 4)
    keys = np.array([int_to_limbs(t, W) for t in terms], dtype=np.uint64).reshape(-1, W)
    coeffs = np.array(list(terms.values()), dtype=np.float64).reshape(-1, 1)
    return OracleSum(W, 1).set(keys, coeffs), W


def _ones(indices, W):
    m = 0
    for q in indices:
        m |= 1 << (2 * (q - 1))
    return np.array(int_to_limbs(m, W), dtype=np.uint64)


@pytest.mark.parametrize("nq,indices,expected", [(4, (2, 4), 1.0), (5, (2,), -1.0), (4, (3,), 1.0), (4, (), 1.0)])
def test_overlapwithcomputational_known_answers(nq, indices, expected):
    # test/test_overlaps.jl:9-46
    for terms, want in (({0: 1.0}, 1.0),
                        ({symboltoint(nq, ["Z", "Z"], [2, 4]): 1.0}, expected),
                        ({symboltoint(nq, ["X", "Z", "Z"], [1, 2, 4]): 1.0}, 0.0),
                        ({symboltoint(nq, ["Y"], [4]): 1.0}, 0.0)):
        s, W = _sum(nq, terms)
        assert s.overlap_kind(2, _ones(indices, W))[0] == want


def test_scalarproduct_known_answers():
    # test/test_overlaps.jl:49-73
    p1 = (symboltoint(4, ["I", "Z"], [1, 2]), 1.0)
    p2 = (symboltoint(4, ["X", "Y"], [2, 3]), -2.0)
    p3 = (symboltoint(4, ["Z", "X"], [3, 4]), 0.2)
    ps, _ = _sum(4, dict([p1, p2]))
    for p, want in ((p3, 0.0), (p1, 1.0), (p2, 4.0)):
        single, _ = _sum(4, dict([p]))
        assert ps.scalar_product(single)[0] == want
        assert single.scalar_product(ps)[0] == want
    assert ps.scalar_product(ps)[0] == 5.0


def test_overlapwithpaulisum_known_answers():
    # test/test_overlaps.jl:83-111: psum = 0.2 I + 0.3 X2 + 0.4 Z3 + 0.5 Z1 Z2 on 3 qubits
    nq = 3
    terms = {0: 0.2, symboltoint(nq, ["X"], [2]): 0.3, symboltoint(nq, ["Z"], [3]): 0.4,
             symboltoint(nq, ["Z", "Z"], [1, 2]): 0.5}
    ps, W = _sum(nq, terms)
    rho, _ = _sum(nq, {0: 1 / 2 ** nq})
    assert rho.scalar_product(ps)[0] * 2 ** nq == pytest.approx(0.2, abs=1e-15)
    assert ps.get_coeff(np.zeros(W, dtype=np.uint64))[0] == 0.2  # overlapwithmaxmixed
    assert rho.overlap_kind(1)[0] == 1 / 2 ** nq                  # overlapwithplus of the maximally mixed state
    assert ps.overlap_kind(0)[0] == pytest.approx(1.1, abs=1e-15)
    # rho = prod_q (I + s_q Z_q) / 2 is |x><x| with x_q = 1 where s_q = -1
    for which in ([0, 0, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1]):
        rho_terms = {0: 1.0}
        for q in range(1, nq + 1):
            s_q = -1.0 if which[q - 1] else 1.0
            nxt = {}
            for t, c in rho_terms.items():
                nxt[t] = nxt.get(t, 0.0) + c / 2
                zt = t | (3 << (2 * (q - 1)))
                nxt[zt] = nxt.get(zt, 0.0) + c * s_q / 2
            rho_terms = nxt
        rh, _ = _sum(nq, rho_terms)
        ones = _ones([q for q in range(1, nq + 1) if which[q - 1]], W)
        assert rh.scalar_product(ps)[0] * 2 ** nq == pytest.approx(ps.overlap_kind(2, ones)[0], abs=1e-15)


@pytest.mark.parametrize("seed", range(4))
def test_readouts_against_dense_matrices(seed):
    rng = np.random.default_rng(seed)
    nq = 5
    terms = {int(t): float(c) for t, c in zip(rng.choice(4 ** nq, 40, replace=False), rng.normal(size=40))}
    terms[0] = 0.3
    other = {int(t): float(c) for t, c in zip(rng.choice(4 ** nq, 60, replace=False), rng.normal(size=60))}
    ps, W = _sum(nq, terms)
    po, _ = _sum(nq, other)
    O = operator_from_terms(terms, nq)
    P = operator_from_terms(other, nq)
    dim = 2 ** nq
    zero = np.zeros(dim)
    zero[0] = 1.0
    plus = np.full(dim, 1 / np.sqrt(dim))
    assert ps.overlap_kind(0)[0] == pytest.approx((zero @ O @ zero).real, abs=1e-12)
    assert ps.overlap_kind(1)[0] == pytest.approx((plus @ O @ plus).real, abs=1e-12)
    for x in rng.choice(dim, 6, replace=False):
        e = np.zeros(dim)
        e[x] = 1.0
        ones = _ones([q for q in range(1, nq + 1) if (x >> (q - 1)) & 1], W)
        assert ps.overlap_kind(2, ones)[0] == pytest.approx((e @ O @ e).real, abs=1e-12)
    assert ps.get_coeff(np.zeros(W, dtype=np.uint64))[0] == pytest.approx(np.trace(O).real / dim, abs=1e-12)
    assert ps.scalar_product(po)[0] == pytest.approx(np.trace(O @ P).real / dim, abs=1e-12)
