"""The oracle's restatement of rewindgradient (src/gradient.jl:26-108), pinned
the way the reference pins its own (test/test_gradient.jl): at
min_abs_coeff = 0 the expectation equals propagate + overlapfunc, the gradient
is not all zero and equals central differences (eps 1e-6, atol 1e-7) of the
dense-pinned forward oracle, for the same kinds of circuits. The sign and the
factor of the commutator dot are fixed by those differences; a wrong one fails
them (test_a_flipped_commutator_sign_fails_the_differences)."""
import math

import numpy as np
import pytest

import pauliprop_hip as pp
from oracle import oracle_adjoint, oracle_propagate, oracle_rewind


def _overlap(s, nq, overlap, onebitinds=(), rho=None):
    keys, coeffs = s.get()
    tot = 0.0
    for k, c in zip(keys, coeffs):
        t = sum(int(k[l]) << (64 * l) for l in range(len(k)))
        codes = [(t >> (2 * q)) & 3 for q in range(nq)]
        if overlap == "zero":
            f = 0 if any(x in (1, 2) for x in codes) else 1
        elif overlap == "plus":
            f = 0 if any(x in (2, 3) for x in codes) else 1
        elif overlap == "computational":
            f = 0 if any(x in (1, 2) for x in codes) else (-1) ** sum(1 for q in onebitinds if codes[q - 1] == 3)
        elif overlap == "maxmixed":
            f = 1 if t == 0 else 0
        else:
            f = (2.0 ** nq) * rho.get(t, 0.0)
        tot += f * float(c[0])
    return tot


def _check(circ, nq, params, obs, overlap="zero", onebitinds=(), rho=None, eps=1e-6, atol=1e-7):
    expec, grad = oracle_rewind(circ, nq, params, obs, min_abs_coeff=0.0, overlap=overlap, onebitinds=onebitinds,
                                rho=rho)
    f = lambda th: _overlap(oracle_propagate(circ, nq, th, obs, min_abs_coeff=0.0)[0], nq, overlap, onebitinds, rho)
    assert expec == pytest.approx(f(np.asarray(params, dtype=float)), abs=1e-12)
    assert len(grad) == len(params)
    assert np.any(grad != 0)
    for k in range(len(params)):
        tp = np.array(params, dtype=float)
        tm = tp.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (f(tp) - f(tm)) / (2 * eps)
        assert abs(fd - grad[k]) < atol, (k, fd, grad[k])
    return expec, grad


R, C, F = pp.PauliRotation, pp.CliffordGate, pp.freeze


def _zz(nq, a, b, c=1.0):
    return {pp.symboltoint(nq, ["Z", "Z"], [a, b]): c}


def test_pure_rotation_circuit():
    circ = [R("X", 1), R("Y", 2), R(["X", "Z"], [1, 3]), R("Z", 4), R(["Y", "Y"], [2, 4])]
    _check(circ, 4, [0.3, -0.7, 1.1, 0.42, -0.15], _zz(4, 1, 3))


def test_interspersed_cliffords():
    circ = [R("X", 1), C("H", 2), R(["Y", "Z"], [1, 3]), C("CNOT", [2, 3]), R("Z", 4), C("S", 4),
            R(["X", "X"], [1, 4])]
    _check(circ, 4, [0.2, -0.5, 0.83, -1.0], _zz(4, 1, 4))


def test_frozen_rotations():
    circ = [R("X", 1), F(R("Y", 2), 0.66), R(["X", "Z"], [1, 3]), F(R("Z", 4), -0.31), R(["Y", "Y"], [2, 4])]
    _check(circ, 4, [0.44, -0.9, 0.12], _zz(4, 1, 3))


def test_cliffords_and_frozen_gates_mixed():
    circ = [R("X", 1), C("H", 2), F(R("Y", 2), 0.5), C("CZ", [2, 3]), R(["X", "Z"], [1, 3]), C("SX", 4),
            F(R("Z", 4), -0.2), R(["Y", "X"], [3, 5]), C("CNOT", [4, 5]), R("X", 5)]
    _check(circ, 5, [0.31, -0.62, 0.9, -0.44], {pp.symboltoint(5, "Z", 3): 1.0})


def _overlap_circuit():
    return [R("X", 1), C("H", 2), R(["Y", "Z"], [1, 3]), F(R("Z", 4), 0.27), C("CNOT", [3, 4]),
            R(["X", "X"], [2, 4])]


@pytest.mark.parametrize("overlap", ["zero", "plus"])
def test_overlap_functions(overlap):
    _check(_overlap_circuit(), 4, [0.18, -0.71, 0.53], _zz(4, 1, 3), overlap=overlap)


def test_computational_overlap():
    _check(_overlap_circuit(), 4, [0.18, -0.71, 0.53], _zz(4, 1, 3), overlap="computational", onebitinds=[1, 3])


def test_maxmixed_gradient_is_zero():
    # unitary propagation conserves the trace, so the maximally mixed overlap is constant
    _, g = oracle_rewind(_overlap_circuit(), 4, [0.18, -0.71, 0.53], _zz(4, 1, 3), min_abs_coeff=0.0,
                         overlap="maxmixed")
    assert np.all(g == 0)


def test_observable_with_several_terms():
    circ = [R("X", 1), C("H", 2), R(["Y", "Z"], [1, 3]), F(R("Z", 4), -0.4), C("CNOT", [3, 4]),
            R(["X", "X"], [2, 4])]
    obs = {pp.symboltoint(4, ["Z", "Z"], [1, 4]): 0.6, pp.symboltoint(4, ["X", "Y"], [2, 3]): -0.3}
    _check(circ, 4, [0.37, -0.28, 0.61], obs)


def test_dense_state_overlap_after_commuting_first_gates():
    # a dense state weighs every string, so a dual term dropped wrongly shows in the gradient
    nq = 4
    circ = [R("Z", 1), C("H", 4), R("X", 2), R("Y", 3), R("Y", 4), R(["X", "X"], [3, 4])]
    rng = np.random.default_rng(3)
    rho = {t: float(rng.normal()) for t in range(256)}
    _check(circ, nq, [0.3, -0.7, 1.1, 0.42, -0.15], _zz(nq, 2, 3), overlap="paulisum", rho=rho)


def test_truncation_to_empty():
    circ = [R("X", 1), R("Z", 2), R("Y", 1)]
    expec, g = oracle_rewind(circ, 2, [0.3, 0.7, 0.2], {pp.symboltoint(2, "Z", 1): 1.0}, min_abs_coeff=10.0)
    assert expec == 0.0 and np.all(g == 0)


@pytest.mark.parametrize("noise", [pp.DepolarizingNoise(2, 0.1), pp.freeze(pp.AmplitudeDampingNoise(2), 0.1)])
def test_noise_is_rejected(noise):
    circ = [R("X", 1), noise, R("Y", 2)]
    with pytest.raises(ValueError):
        oracle_rewind(circ, 3, [0.3, 0.4], _zz(3, 1, 2))


@pytest.mark.parametrize("trial", range(6))
def test_random_small_circuits(trial):
    rng = np.random.default_rng(33 + trial)
    nq = 5
    circ = []
    for _ in range(8):
        u = rng.random()
        if u < 0.3:
            circ.append(C(["H", "S", "SX"][rng.integers(0, 3)], int(rng.integers(1, nq + 1))))
        elif u < 0.45:
            a, b = rng.permutation(nq)[:2] + 1
            circ.append(C(["CNOT", "CZ"][rng.integers(0, 2)], [int(a), int(b)]))
        elif u < 0.6:
            if rng.random() < 0.5:
                circ.append(F(R("XYZ"[rng.integers(0, 3)], int(rng.integers(1, nq + 1))), float(rng.normal())))
            else:
                a, b = rng.permutation(nq)[:2] + 1
                circ.append(F(R([str("XYZ"[rng.integers(0, 3)]) for _ in range(2)], [int(a), int(b)]),
                              float(rng.normal())))
        else:
            if rng.random() < 0.5:
                circ.append(R("XYZ"[rng.integers(0, 3)], int(rng.integers(1, nq + 1))))
            else:
                a, b = rng.permutation(nq)[:2] + 1
                circ.append(R([str("XYZ"[rng.integers(0, 3)]) for _ in range(2)], [int(a), int(b)]))
    P = pp.countparameters(circ)
    if P == 0:
        circ.append(R("X", 1))
        P = 1
    params = rng.normal(size=P)
    a, b = rng.permutation(nq)[:2] + 1
    obs = _zz(nq, int(a), int(b))
    expec, grad = oracle_rewind(circ, nq, params, obs, min_abs_coeff=0.0)
    f = lambda th: oracle_propagate(circ, nq, th, obs, min_abs_coeff=0.0)[0].overlap_zero()[0]
    assert expec == pytest.approx(f(params), abs=1e-12)
    for k in range(P):
        tp, tm = params.copy(), params.copy()
        tp[k] += 1e-6
        tm[k] -= 1e-6
        assert abs((f(tp) - f(tm)) / 2e-6 - grad[k]) < 1e-7


def test_equals_the_fixed_mask_adjoint_without_truncation():
    # with nothing truncated both gradients are the exact derivative
    rng = np.random.default_rng(5)
    n = 6
    circ = []
    for _ in range(3):
        pp.rylayer(circ, n)
        pp.rzlayer(circ, n)
        pp.rzzlayer(circ, pp.staircasetopology(n))
    P = pp.countparameters(circ)
    th = rng.normal(0, 0.3, size=P)
    obs = {pp.symboltoint(n, ["Z", "Z"], [i, i + 1]): 1.0 for i in range(1, n)}
    l1, g1 = oracle_rewind(circ, n, th, obs, min_abs_coeff=0.0)
    l2, g2 = oracle_adjoint(circ, n, th, obs, min_abs_coeff=0.0)
    assert l1 == pytest.approx(l2[0], abs=1e-13)
    np.testing.assert_allclose(g1, g2[:, 0], atol=1e-12, rtol=0)


def test_a_flipped_commutator_sign_fails_the_differences():
    # the gradient's sign is fixed by the differences: its negative is far from them
    circ = [R("X", 1), R("Y", 2), R(["X", "Z"], [1, 3]), R("Z", 4), R(["Y", "Y"], [2, 4])]
    params = np.array([0.3, -0.7, 1.1, 0.42, -0.15])
    obs = _zz(4, 1, 3)
    _, g = oracle_rewind(circ, 4, params, obs, min_abs_coeff=0.0)
    f = lambda th: oracle_propagate(circ, 4, th, obs, min_abs_coeff=0.0)[0].overlap_zero()[0]
    fd = np.array([(f(params + 1e-6 * e) - f(params - 1e-6 * e)) / 2e-6 for e in np.eye(5)])
    assert np.max(np.abs(-g - fd)) > 1e-2
    assert np.max(np.abs(g - fd)) < 1e-7


@pytest.mark.parametrize("overlap,ones", [("zero", ()), ("plus", ()), ("computational", (1, 3)), ("maxmixed", ())])
def test_fixed_mask_adjoint_with_cliffords_equals_rewind_without_truncation(overlap, ones):
    circ = _overlap_circuit() + [C("SY", 1), R(["Z", "X"], [2, 3]), C("ZZPIHALF", [1, 4]), R("Y", 4)]
    P = pp.countparameters(circ)
    params = np.linspace(-0.9, 0.8, P)
    obs = {pp.symboltoint(4, ["Z", "Z"], [1, 3]): 1.0, pp.symboltoint(4, "X", 2): 0.4}
    l1, g1 = oracle_rewind(circ, 4, params, obs, min_abs_coeff=0.0, overlap=overlap, onebitinds=ones)
    l2, g2 = oracle_adjoint(circ, 4, params, obs, min_abs_coeff=0.0, overlap=overlap, onebitinds=ones)
    assert l1 == pytest.approx(l2[0], abs=1e-13)
    np.testing.assert_allclose(g1, g2[:, 0], atol=1e-12, rtol=0)
    if overlap != "maxmixed":
        _check(circ, 4, params, obs, overlap=overlap, onebitinds=ones)
