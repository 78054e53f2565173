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


# ------------------------------------------------------------------ Clifford tables and amplitude damping
# The Clifford tables (src/Gates/cliffordgates.jl:47-59) and the amplitude-damping
# rule (src/Propagation/specializations.jl:258-276) checked against the gates'
# matrices: U^dag P U for every Pauli P, and sum_k K_k^dag P K_k.

BASE_CLIFFORDS = ["H", "X", "Y", "Z", "SX", "SY", "S", "SZ", "CNOT", "CZ", "SWAP", "ZZPIHALF"]


def _single(expansion):
    assert len(expansion) == 1, expansion
    (q, v), = expansion.items()
    assert abs(abs(v) - 1.0) < 1e-12, expansion
    return q, int(round(v))


@pytest.mark.parametrize("symbol", BASE_CLIFFORDS)
def test_clifford_table_is_the_conjugation(symbol):
    table = pp.CLIFFORD_MAP[symbol]
    nq = 1 if len(table) == 4 else 2
    U = dense.embed(dense.CLIFFORD_UNITARIES[symbol], list(range(1, nq + 1)), nq)
    transposed = pp.transposecliffordmap(table)
    for p in range(4 ** nq):
        P = dense.pauli_matrix(dense.term_codes(p, nq), nq)
        # Heisenberg action (the table itself)
        assert _single(dense.pauli_expand(U.conj().T @ P @ U, nq, tol=1e-12)) == (table[p][0], table[p][1]), \
            (symbol, p)
        # Schrodinger action (the transposed table, src/Gates/cliffordgates.jl:94-98)
        assert _single(dense.pauli_expand(U @ P @ U.conj().T, nq, tol=1e-12)) == (transposed[p][0], transposed[p][1]), \
            (symbol, p)


def test_clifford_table_check_catches_a_flipped_sign():
    # a wrong but self-consistent entry (sign of S on X) is caught by the conjugation check
    table = list(pp.CLIFFORD_MAP["S"])
    table[1] = (table[1][0], -table[1][1])
    U = dense.CLIFFORD_UNITARIES["S"]
    P = dense.pauli_matrix({1: 1}, 1)
    assert _single(dense.pauli_expand(U.conj().T @ P @ U, 1, tol=1e-12)) != table[1]
    assert pp.transposecliffordmap(pp.transposecliffordmap(table)) == table  # the old check passes it


@pytest.mark.parametrize("n", [1, 2, 3])
@pytest.mark.parametrize("gamma", [0.0, 0.3, 1.0])
def test_amplitude_damping_against_kraus_channel(n, gamma):
    from oracle import OracleSum
    from pauliprop_hip.paulis import int_to_limbs
    rng = np.random.default_rng(100 * n + int(10 * gamma))
    terms = list(range(4 ** n))
    coeffs = rng.uniform(-1, 1, size=len(terms))
    for q in range(1, n + 1):
        s = OracleSum(1, 1)
        s.set(np.array([int_to_limbs(t, 1) for t in terms], dtype=np.uint64), coeffs)
        s.amplitude_damping(q - 1, gamma, max_weight=math.inf, min_abs=0.0)
        op = dense.heisenberg_step_damping(dense.operator_from_terms(dict(zip(terms, coeffs)), n), q, gamma, n)
        want = dense.pauli_expand(op, n, tol=0.0)
        keys, c = s.get()
        got = {int(k[0]): float(v[0]) for k, v in zip(keys, c)}
        for t in set(want) | set(got):
            assert abs(got.get(t, 0.0) - want.get(t, 0.0)) < 1e-12, (n, q, gamma, t)


def test_amplitude_damping_closed_form():
    # I -> I, X -> sqrt(1-g) X, Y -> sqrt(1-g) Y, Z -> (1-g) Z + g I, from the Kraus operators
    g = 0.37
    for code in range(4):
        op = dense.heisenberg_step_damping(dense.pauli_matrix({1: code} if code else {}, 1), 1, g, 1)
        got = dense.pauli_expand(op, 1, tol=1e-15)
        want = {0: {0: 1.0}, 1: {1: math.sqrt(1 - g)}, 2: {2: math.sqrt(1 - g)}, 3: {3: 1 - g, 0: g}}[code]
        assert set(got) == set(want) and all(abs(got[k] - want[k]) < 1e-14 for k in want), (code, got)


def _mixed_circuit(n, depth, rng):
    circ = []
    syms1 = ["H", "X", "Y", "Z", "SX", "SY", "S"]
    syms2 = ["CNOT", "CZ", "SWAP", "ZZPIHALF"]
    for _ in range(depth):
        kind = rng.integers(0, 6)
        q = int(rng.integers(1, n + 1))
        if kind == 0:
            circ.append(pp.PauliRotation(str("XYZ"[rng.integers(0, 3)]), q))
        elif kind == 1 and n >= 2:
            a, b = rng.choice(np.arange(1, n + 1), size=2, replace=False)
            circ.append(pp.PauliRotation([str("XYZ"[rng.integers(0, 3)]), str("XYZ"[rng.integers(0, 3)])],
                                         [int(a), int(b)]))
        elif kind == 2:
            circ.append(pp.CliffordGate(syms1[rng.integers(0, len(syms1))], q))
        elif kind == 3 and n >= 2:
            a, b = rng.choice(np.arange(1, n + 1), size=2, replace=False)
            circ.append(pp.CliffordGate(syms2[rng.integers(0, len(syms2))], [int(a), int(b)]))
        elif kind == 4:
            circ.append(pp.freeze(pp.AmplitudeDampingNoise(q), float(rng.uniform(0, 0.4))))
        else:
            circ.append(pp.DepolarizingNoise(q, float(rng.uniform(0, 0.2))))
    return circ


def _dense_mixed(hcirc, hth):
    out, k = [], 0
    for g in hcirc:
        inner = g.gate if isinstance(g, pp.FrozenGate) else g
        if isinstance(inner, pp.PauliRotation):
            th = g.parameter if isinstance(g, pp.FrozenGate) else hth[k]
            if not isinstance(g, pp.FrozenGate):
                k += 1
            out.append(("rot", {q: pp.paulis.SYMBOLS[s] for s, q in zip(inner.symbols, inner.qinds)}, th))
        elif isinstance(inner, pp.CliffordGate):
            out.append(("clifford", inner.symbol, inner.qinds))
        elif isinstance(inner, pp.AmplitudeDampingNoise):
            out.append(("damp", inner.qind, g.parameter))
        else:
            out.append(("noise", inner.qind, set(inner.damped), g.parameter))
    return out


@pytest.mark.parametrize("seed", range(10))
def test_mixed_circuit_with_cliffords_and_damping_against_dense(seed):
    rng = np.random.default_rng(500 + seed)
    n = int(rng.integers(2, 6))
    circ = _mixed_circuit(n, 18, rng)
    P = pp.countparameters(circ)
    th = rng.uniform(-np.pi, np.pi, size=P)
    obs = {pp.symboltoint(n, "Z", 1): 1.0, pp.symboltoint(n, ["X", "Y"], [1, n]): 0.5,
           pp.symboltoint(n, "Y", n): -0.25}
    s, _ = oracle_propagate(circ, n, th, obs, min_abs_coeff=0.0)
    hc, hth = pp.toheisenberg(circ, list(th))
    op = dense.dense_propagate(_dense_mixed(hc, hth), n, obs)
    want = dense.pauli_expand(op, n, tol=1e-13)
    keys, coeffs = s.get()
    got = {}
    for k, c in zip(keys, coeffs):
        got[int(k[0])] = got.get(int(k[0]), 0.0) + float(c[0])
    for t in set(want) | set(got):
        assert abs(got.get(t, 0.0) - want.get(t, 0.0)) < 1e-12, (seed, t, got.get(t), want.get(t))
    assert s.overlap_zero()[0] == pytest.approx(dense.zero_state_expectation(op), abs=1e-12)
