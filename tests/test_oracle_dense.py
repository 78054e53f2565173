"""Whole propagations of the oracle against dense brute force (numpy, 2^n x
2^n matrices) for n <= 5 qubits, no truncation: term sets, coefficients and
expectation values. The dense side knows only the Pauli matrices and the
gate definitions."""
import math

import numpy as np
import pytest

import dense
from oracle import oracle_propagate, oracle_adjoint
import pauliprop_hip as pp


def _random_circuit(n, depth, rng, noise):
    circ = []
    for _ in range(depth):
        kind = rng.integers(0, 4 if noise else 3)
        if kind == 0:
            q = int(rng.integers(1, n + 1))
            circ.append(pp.PauliRotation(str("XYZ"[rng.integers(0, 3)]), q))
        elif kind == 1:
            a, b = rng.choice(np.arange(1, n + 1), size=2, replace=False)
            s = "XYZ"[rng.integers(0, 3)]
            circ.append(pp.PauliRotation([s, s], [int(a), int(b)]))
        elif kind == 2:
            k = int(rng.integers(1, n + 1))
            qs = rng.choice(np.arange(1, n + 1), size=k, replace=False)
            circ.append(pp.PauliRotation([str("XYZ"[rng.integers(0, 3)]) for _ in qs], [int(q) for q in qs]))
        else:
            q = int(rng.integers(1, n + 1))
            maker = [pp.DepolarizingNoise, pp.PauliXNoise, pp.PauliZNoise, pp.PauliYDamping][rng.integers(0, 4)]
            circ.append(maker(q, float(rng.uniform(0, 0.3))))
    return circ


def _dense_gates(hcirc, hth, n):
    out = []
    k = 0
    for g in hcirc:
        if isinstance(g, pp.PauliRotation):
            out.append(("rot", {q: pp.paulis.SYMBOLS[s] for s, q in zip(g.symbols, g.qinds)}, hth[k]))
            k += 1
        else:
            inner = g.gate
            out.append(("noise", inner.qind, set(inner.damped), g.parameter))
    return out


@pytest.mark.parametrize("seed", range(12))
def test_random_circuit_against_dense(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 6))
    circ = _random_circuit(n, 14, rng, noise=seed % 2 == 1)
    P = pp.countparameters(circ)
    th = rng.uniform(-np.pi, np.pi, size=P)
    obs = {pp.symboltoint(n, "Z", 1): 1.0, pp.symboltoint(n, ["X", "Y"], [1, n]): 0.5}
    s, _ = oracle_propagate(circ, n, th, obs, min_abs_coeff=0.0)
    hc, hth = pp.toheisenberg(circ, list(th))
    op = dense.dense_propagate(_dense_gates(hc, hth, n), n, obs)
    want = dense.pauli_expand(op, n, tol=1e-12)
    keys, coeffs = s.get()
    got = {int(k[0]): float(c[0]) for k, c in zip(keys, coeffs)}
    for t in set(want) | set(got):
        assert abs(got.get(t, 0.0) - want.get(t, 0.0)) < 1e-12, (t, got.get(t), want.get(t))
    # terms the oracle keeps but dense drops are exact cancellations only
    for t in set(got) - set(want):
        assert abs(got[t]) < 1e-12
    assert s.overlap_zero()[0] == pytest.approx(dense.zero_state_expectation(op), abs=1e-12)


@pytest.mark.parametrize("noise", [False, True])
def test_adjoint_against_central_differences(noise):
    # test/test_gradient.jl:9-31: central differences, eps 1e-6, atol 1e-7, no truncation
    rng = np.random.default_rng(11 + noise)
    n = 6
    circ = []
    for _ in range(3):
        pp.rylayer(circ, n)
        pp.rzlayer(circ, n)
        pp.rzzlayer(circ, pp.staircasetopology(n))
        if noise:
            pp.depolarizinglayer(circ, n, 0.02)
    P = pp.countparameters(circ)
    th = rng.normal(0, 0.3, size=P)
    obs = {pp.symboltoint(n, ["Z", "Z"], [i, i + 1]): 1.0 for i in range(1, n)}
    loss, grad = oracle_adjoint(circ, n, th, obs, min_abs_coeff=0.0)
    f = lambda t: oracle_propagate(circ, n, t, obs, min_abs_coeff=0.0)[0].overlap_zero()[0]
    assert loss[0] == pytest.approx(f(th), abs=1e-14)
    eps = 1e-6
    for k in range(P):
        tp, tm = th.copy(), th.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (f(tp) - f(tm)) / (2 * eps)
        assert abs(fd - grad[k, 0]) < 1e-7, (k, fd, grad[k, 0])


def test_batched_adjoint_columns_equal_single_runs():
    rng = np.random.default_rng(2)
    n = 5
    circ = []
    for _ in range(2):
        pp.rxlayer(circ, n)
        pp.rzzlayer(circ, pp.bricklayertopology(n))
        pp.depolarizinglayer(circ, n, 0.05)
    P = pp.countparameters(circ)
    B = 3
    th = rng.uniform(-1, 1, size=(P, B))
    obs = {pp.symboltoint(n, "Z", 3): 1.0}
    loss, grad = oracle_adjoint(circ, n, th, obs, min_abs_coeff=0.0, batch=B)
    for b in range(B):
        l1, g1 = oracle_adjoint(circ, n, th[:, b], obs, min_abs_coeff=0.0)
        assert loss[b] == pytest.approx(l1[0], abs=1e-15)
        np.testing.assert_allclose(grad[:, b], g1[:, 0], atol=1e-15)
