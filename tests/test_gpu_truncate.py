"""GPU parity of maxabscoeff, truncate! and propagate with min_rel_coeff
(src/Base/Primitives/maxabscoeff.jl:7, src/Propagation/generics.jl:192-250,
src/Propagation/specializations.jl:367-377): bit for bit against the oracle
for single sums of every key width and for whole circuits of rotations,
Cliffords, Pauli noise and amplitude damping, plus the reference's known
answer (test/test_truncations.jl:122-129)."""
import math

import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import OracleSum, oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.paulis import int_to_limbs, symboltoint

pytestmark = pytest.mark.gpu


def _pair(nq, n, B, rng):
    W = pp.key_words(nq)
    terms = set()
    while len(terms) < n:
        t = 0
        for q in rng.choice(nq, size=int(rng.integers(0, min(nq, 9) + 1)), replace=False):
            t |= int(rng.integers(1, 4)) << (2 * int(q))
        terms.add(t)
    keys = np.array([int_to_limbs(t, W) for t in sorted(terms)], dtype=np.uint64).reshape(-1, W)
    c = rng.normal(size=(n, B)) * np.exp(2.0 * rng.normal(size=(n, B)))
    dev = pp.HIPPauliSum(nq, batch=B, capacity=2 * n)
    dev.upload(keys, c)
    return dev, OracleSum(W, B).set(keys, c)


def test_known_answer_on_device():
    print("truncate known answer", flush=True)
    s = pp.HIPPauliSum.from_paulisum(pp.PauliSum(1, {1: 1.0, 2: 0.1}))
    assert s.maxabscoeff() == 1.0
    assert s.truncate(min_abs_coeff=0.0, min_rel_coeff=0.2) == 0.2
    assert len(s) == 1 and s.getcoeff(1) == 1.0
    e = pp.HIPPauliSum(30)
    assert e.maxabscoeff() == 0.0
    e.truncate(min_rel_coeff=0.5)
    assert len(e) == 0
    b = pp.HIPPauliSum(30, batch=2)
    with pytest.raises(pp._lib.PPError):
        b.truncate(min_rel_coeff=0.1)


@pytest.mark.parametrize("nq,B", [(20, 1), (60, 1), (120, 1), (60, 4)])
def test_maxabs_and_truncate_match_oracle(nq, B):
    print(f"truncate nq={nq} B={B}", flush=True)
    rng = np.random.default_rng(nq + B)
    for n in (1, 777, 300000):
        dev, ref = _pair(nq, n, B, rng)
        got = dev.maxabscoeff()
        np.testing.assert_array_equal(np.atleast_1d(got), ref.max_abs())
        rels = (None, 1e-3, 0.2) if B == 1 else (None,)
        for mw, mac, rel in ((math.inf, 1e-2, None), (5, 0.0, None), (math.inf, 0.0, 1e-3), (7, 1e-4, 0.2)):
            if rel not in rels:
                continue
            dev, ref = _pair(nq, n, B, np.random.default_rng(n))
            t_dev = dev.truncate(mw, mac, rel)
            t_ref = ref.truncate(mw, mac, rel)
            assert t_dev == t_ref
            assert_bitwise(*dev.download(), *ref.get(), f"truncate n={n} mw={mw} mac={mac} rel={rel}")


@pytest.mark.parametrize("nq", [16, 70])
@pytest.mark.parametrize("rel", [0.0, 1e-4, 1e-2])
def test_propagate_with_relative_threshold_matches_oracle(nq, rel):
    print(f"propagate min_rel_coeff={rel} nq={nq}", flush=True)
    rng = np.random.default_rng(int(rel * 1e4) + nq)
    circ = []
    for layer in range(3):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, pp.bricklayertopology(nq))
        for q in range(1, nq, 5):
            circ.append(pp.CliffordGate("CNOT", [q, q + 1]))
        pp.depolarizinglayer(circ, nq, 0.01)
        circ.append(pp.AmplitudeDampingNoise(layer + 2))
        pp.rylayer(circ, nq)
    th = np.array([rng.uniform(0.05, 0.2) if isinstance(g, pp.AmplitudeDampingNoise) else rng.uniform(-0.7, 0.7)
                   for g in circ if g.parametrized])
    obs = pp.PauliSum(nq, {symboltoint(nq, ["Z"], [nq // 2]): 1.0,
                           symboltoint(nq, ["X", "Y"], [3, 4]): -0.4})
    for mw, mac in ((math.inf, 1e-6), (5, 1e-5)):
        dev = pp.propagate(circ, obs, th, max_weight=mw, min_abs_coeff=mac, min_rel_coeff=rel)
        ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=mw, min_abs_coeff=mac, min_rel_coeff=rel)
        assert_bitwise(*dev.download(), *ref.get(), f"min_rel_coeff={rel} max_weight={mw}")
        if rel == 0.0:
            plain = pp.propagate(circ, obs, th, max_weight=mw, min_abs_coeff=mac)
            assert_bitwise(*dev.download(), *plain.download(), "min_rel_coeff=0 vs none")


def test_schrodinger_circuit_with_relative_threshold_matches_oracle():
    # heisenberg=False: gates in circuit order, imaginary rotations normalized,
    # every gate but a Clifford followed by truncate! with min_rel_coeff
    print("Schrodinger circuit with min_rel_coeff", flush=True)
    nq = 12
    circ = []
    for layer in range(3):
        pp.rxlayer(circ, nq)
        for q in (2, 6, 9):
            circ.append(pp.ImaginaryPauliRotation(["Z", "Z"], [q, q + 1]))
        circ.append(pp.CliffordGate("CNOT", [layer + 3, layer + 4]))
        pp.rzzlayer(circ, pp.bricklayertopology(nq))
        pp.depolarizinglayer(circ, nq, 0.02)
    th = np.random.default_rng(31).uniform(-0.5, 0.5, size=pp.countparameters(circ))
    rho = pp.PauliSum(nq, {0: 1.0, symboltoint(nq, ["Z"], [4]): 0.3})
    for rel in (1e-5, 1e-4, 1e-3):
        dev = pp.propagate(circ, rho, th, max_weight=6, min_abs_coeff=1e-7, min_rel_coeff=rel, heisenberg=False)
        ref, _ = oracle_propagate(circ, nq, th, rho.terms, max_weight=6, min_abs_coeff=1e-7, min_rel_coeff=rel,
                                  heisenberg=False)
        assert_bitwise(*dev.download(), *ref.get(), f"Schrodinger min_rel_coeff={rel}")
        assert len(dev) > 100
