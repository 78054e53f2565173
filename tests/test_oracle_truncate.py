"""The oracle's truncate!/maxabscoeff and min_rel_coeff propagation against the
reference's own tests (test/test_truncations.jl:90-129): buildtruncfunc's
known answer, min_rel_coeff = 0 adds nothing beyond min_abs_coeff, a larger
min_rel_coeff truncates more; plus what the rule fixes: after a gate that is
not a Clifford no kept coefficient lies below min_rel_coeff times the largest,
and a gate-by-gate walk of untruncated gates with a plain filter gives the
same sum."""
import math

import numpy as np
import pytest

from oracle import OracleSum, oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.paulis import countweight, limbs_to_int, symboltoint


def _brick(nq=10, layers=3, seed=42):
    circ = pp.hardwareefficientcircuit(nq, layers)
    th = np.random.default_rng(seed).normal(size=pp.countparameters(circ))
    return circ, th


def _terms(s):
    k, c = s.get()
    return {limbs_to_int(a): float(b[0]) for a, b in zip(k, c)}


def test_buildtruncfunc_known_answer():
    # test/test_truncations.jl:122-129: X at 1.0 stays, Y at 0.1 goes under 0.2 * 1.0
    s = OracleSum(1, 1).set(np.array([[1], [2]], dtype=np.uint64), np.array([[1.0], [0.1]]))
    assert s.max_abs()[0] == 1.0
    assert s.truncate(math.inf, 0.0, 0.2) == 0.2
    assert _terms(s) == {1: 1.0}


def test_max_abs_of_columns_and_of_an_empty_sum():
    rng = np.random.default_rng(3)
    c = rng.normal(size=(50, 3))
    s = OracleSum(1, 3).set(np.arange(1, 51, dtype=np.uint64).reshape(-1, 1), c)
    np.testing.assert_array_equal(s.max_abs(), np.abs(c).max(axis=0))
    assert list(OracleSum(2, 2).max_abs()) == [0.0, 0.0]


def test_zero_relative_threshold_adds_nothing():
    # test/test_truncations.jl:97-101
    circ, th = _brick()
    obs = {symboltoint(10, ["Z"], [5]): 1.0}
    a, _ = oracle_propagate(circ, 10, th, obs, min_abs_coeff=1e-8)
    b, _ = oracle_propagate(circ, 10, th, obs, min_abs_coeff=1e-8, min_rel_coeff=0.0)
    ka, ca = a.get()
    kb, cb = b.get()
    np.testing.assert_array_equal(ka, kb)
    np.testing.assert_array_equal(ca, cb)


def test_larger_relative_threshold_truncates_more():
    # test/test_truncations.jl:103-116
    circ, th = _brick()
    obs = {symboltoint(10, ["Z"], [5]): 1.0}
    prev, _ = oracle_propagate(circ, 10, th, obs, min_abs_coeff=1e-8)
    prev = full = len(prev)
    for rel in (1e-4, 1e-3, 1e-2):
        s, _ = oracle_propagate(circ, 10, th, obs, min_abs_coeff=1e-8, min_rel_coeff=rel)
        assert len(s) <= prev
        prev = len(s)
        t = _terms(s)
        m = max(abs(v) for v in t.values())
        # the circuit ends on a rotation: its truncate! kept nothing below rel * max
        assert min(abs(v) for v in t.values()) >= rel * m
    assert prev < full // 2


@pytest.mark.parametrize("seed", range(3))
def test_gate_by_gate_walk_with_a_plain_filter(seed):
    # each gate alone without truncation (a Clifford with it), then the
    # threshold max(rel * max|c|, min_abs) and the weight cap applied in Python
    rng = np.random.default_rng(seed)
    nq = 6
    circ = [pp.PauliRotation("X", q) for q in range(1, nq + 1)]
    circ += [pp.CliffordGate("CNOT", [2, 3]), pp.DepolarizingNoise(4)]
    circ += [pp.PauliRotation(["Z", "Z"], [q, q + 1]) for q in range(1, nq)]
    circ += [pp.AmplitudeDampingNoise(2), pp.CliffordGate("H", [5])]
    th = np.array([rng.uniform(-1, 1) if isinstance(g, pp.PauliRotation) else rng.uniform(0.05, 0.3)
                   for g in circ if g.parametrized])
    obs = {symboltoint(nq, ["Y", "Z"], [3, 4]): 1.0, symboltoint(nq, ["X"], [1]): 0.3}
    rel, mac, mw = 0.05, 1e-6, 4
    s, _ = oracle_propagate(circ, nq, th, obs, max_weight=mw, min_abs_coeff=mac, min_rel_coeff=rel)
    hc, hth = pp.toheisenberg(circ, list(th))
    terms, k = dict(obs), 0
    for gate in hc:
        n = pp.countparameters([gate])
        clifford = isinstance(gate, pp.CliffordGate)
        step, _ = oracle_propagate([gate], nq, hth[k:k + n], terms, max_weight=mw if clifford else math.inf,
                                   min_abs_coeff=mac if clifford else 0.0)
        k += n
        terms = _terms(step)
        if not clifford:
            thr = max(rel * max((abs(v) for v in terms.values()), default=0.0), mac)
            terms = {t: v for t, v in terms.items() if abs(v) >= thr and countweight(t) <= mw}
    assert _terms(s) == terms
    assert len(terms) > 3
