"""The oracle's ImaginaryPauliRotation and Schrodinger-picture propagation,
pinned by the reference's known answers (test/test_imaginary.jl: the two-term
result and its normalization, the closed-form thermal energies of a
product-state X Hamiltonian, cooling of a diagonal Ising Hamiltonian) and by
dense matrices (e^{-tau P/2} rho e^{-tau P/2} and U rho U^dag)."""
import math

import numpy as np
import pytest

from oracle import oracle_propagate
from dense import operator_from_terms, pauli_expand, pauli_matrix, rotation_unitary, term_codes
import pauliprop_hip as pp
from pauliprop_hip.paulis import symboltoint


def _terms(s):
    keys, coeffs = s.get()
    from pauliprop_hip.paulis import limbs_to_int
    return {limbs_to_int(k): float(c[0]) for k, c in zip(keys, coeffs)}


def _scalar(a: dict, b: dict) -> float:
    return math.fsum(c * b[t] for t, c in a.items() if t in b)


@pytest.mark.parametrize("nq", [5, 40, 127])
def test_imaginary_rotation_known_answers(nq):
    # test/test_imaginary.jl:1-57: the identity evolves into cosh I - sinh P,
    # normalized by the identity's coefficient unless normalize_coeffs=False
    rng = np.random.default_rng(nq)
    for support in (1, 2, 3, 4):
        symbols = [str(s) for s in rng.choice(["X", "Y", "Z"], support)]
        qinds = [int(q) for q in rng.choice(np.arange(1, nq + 1), support, replace=False)]
        gate = pp.ImaginaryPauliRotation(symbols, qinds)
        P = symboltoint(nq, symbols, qinds)
        tau = float(rng.uniform())
        s, _ = oracle_propagate([gate], nq, [tau], {0: 1.0}, min_abs_coeff=0, heisenberg=False)
        t = _terms(s)
        assert len(t) == 2
        assert t[0] == pytest.approx(1.0, rel=1e-15)
        assert t[P] == pytest.approx(-math.sinh(tau) / math.cosh(tau), rel=1e-14)
        s, _ = oracle_propagate([gate], nq, [tau], {0: 1.0}, min_abs_coeff=0, heisenberg=False,
                                normalize_coeffs=False)
        t = _terms(s)
        assert len(t) == 2 and t[0] == math.cosh(tau) and t[P] == -math.sinh(tau)
        with pytest.raises(RuntimeError):
            oracle_propagate([gate], nq, [tau], {0: 1.0})


@pytest.mark.parametrize("seed", range(4))
def test_imaginary_rotation_against_dense(seed):
    rng = np.random.default_rng(100 + seed)
    n = 5
    rho = {int(t): float(c) for t, c in zip(rng.choice(4 ** n, 24, replace=False), rng.normal(size=24))}
    rho[0] = 1.0
    support = int(rng.integers(1, 4))
    qinds = [int(q) for q in rng.choice(np.arange(1, n + 1), support, replace=False)]
    symbols = [str(s) for s in rng.choice(["X", "Y", "Z"], support)]
    tau = float(rng.uniform(0.1, 1.5))
    P = pauli_matrix(term_codes(symboltoint(n, symbols, qinds), n), n)
    E = math.cosh(tau / 2) * np.eye(2 ** n) - math.sinh(tau / 2) * P  # e^{-tau P / 2}
    out = E @ operator_from_terms(rho, n) @ E
    want = pauli_expand(out, n, tol=1e-14)
    for normalize in (True, False):
        s, _ = oracle_propagate([pp.ImaginaryPauliRotation(symbols, qinds)], n, [tau], rho, min_abs_coeff=0,
                                heisenberg=False, normalize_coeffs=normalize)
        got = _terms(s)
        scale = want[0] if normalize else 1.0
        for t in set(got) | set(want):
            assert got.get(t, 0.0) == pytest.approx(want.get(t, 0.0) / scale, abs=1e-12), t


@pytest.mark.parametrize("seed", range(3))
def test_schrodinger_rotations_against_dense(seed):
    # heisenberg=False: gates in circuit order, rho -> U rho U^dag
    rng = np.random.default_rng(200 + seed)
    n = 5
    rho = {int(t): float(c) for t, c in zip(rng.choice(4 ** n, 12, replace=False), rng.normal(size=12))}
    circ, dense_gates = [], []
    for _ in range(8):
        support = int(rng.integers(1, 3))
        qinds = [int(q) for q in rng.choice(np.arange(1, n + 1), support, replace=False)]
        symbols = [str(s) for s in rng.choice(["X", "Y", "Z"], support)]
        circ.append(pp.PauliRotation(symbols, qinds))
        dense_gates.append(term_codes(symboltoint(n, symbols, qinds), n))
    th = rng.uniform(-1, 1, size=len(circ))
    op = operator_from_terms(rho, n)
    for codes, theta in zip(dense_gates, th):
        U = rotation_unitary(codes, theta, n)
        op = U @ op @ U.conj().T
    want = pauli_expand(op, n, tol=1e-14)
    s, _ = oracle_propagate(circ, n, th, rho, min_abs_coeff=0, heisenberg=False)
    got = _terms(s)
    for t in set(got) | set(want):
        assert got.get(t, 0.0) == pytest.approx(want.get(t, 0.0), abs=1e-12), t


def test_transposed_clifford_maps_invert_the_maps():
    # src/Gates/cliffordgates.jl:94-98: the transpose undoes the map, signs included
    for sym, table in list(pp.CLIFFORD_MAP.items()):
        if sym.endswith("_TRANSPOSE"):
            continue
        tt = pp.transposecliffordmap(table)
        assert pp.transposecliffordmap(tt) == table
        for p, (q, s) in enumerate(table):
            assert tt[q] == (p, s)


def test_cooling_product_state_x_hamiltonian_closed_form():
    # test/test_imaginary.jl:91-131: thermal energies -sum h tanh(beta h) to 1e-9
    nq = 4
    h = [0.8, -0.5, 1.2, -0.3]
    H = {symboltoint(nq, ["X"], [i + 1]): h[i] for i in range(nq)}
    circ = [pp.ImaginaryPauliRotation("X", i + 1) for i in range(nq)]
    rho = {0: 1.0}
    energies = [_scalar(H, rho)]
    for layer in range(1, 11):
        s, _ = oracle_propagate(circ, nq, list(h), rho, min_abs_coeff=0, heisenberg=False)
        rho = _terms(s)
        energies.append(_scalar(H, rho))
        exact = -sum(hi * math.tanh(layer * hi) for hi in h)
        assert energies[-1] == pytest.approx(exact, abs=1e-9)
    assert energies == sorted(energies, reverse=True)
    assert energies[-1] == pytest.approx(-sum(abs(x) for x in h), abs=1e-2)


def test_cooling_diagonal_ising_hamiltonian():
    # test/test_imaginary.jl:35-88
    nq = 4
    hs = [0.6, -0.4, 0.9, -0.7]
    pairs = [(1, 2), (2, 3), (3, 4)]
    Js = [0.5, -0.6, 0.4]
    H = {symboltoint(nq, ["Z"], [i + 1]): hs[i] for i in range(nq)}
    for (a, b), J in zip(pairs, Js):
        H[symboltoint(nq, ["Z", "Z"], [a, b])] = J
    ground = min(sum(hs[i] * z[i] for i in range(nq)) + sum(J * z[a - 1] * z[b - 1] for (a, b), J in zip(pairs, Js))
                 for z in ([1 - 2 * ((s >> i) & 1) for i in range(nq)] for s in range(2 ** nq)))
    circ = [pp.ImaginaryPauliRotation("Z", i + 1) for i in range(nq)]
    circ += [pp.ImaginaryPauliRotation(["Z", "Z"], [a, b]) for a, b in pairs]
    params = list(hs) + list(Js)
    rho = {0: 1.0}
    energies = [_scalar(H, rho)]
    for _ in range(10):
        s, _ = oracle_propagate(circ, nq, params, rho, min_abs_coeff=0, heisenberg=False)
        rho = _terms(s)
        energies.append(_scalar(H, rho))
    assert energies == sorted(energies, reverse=True)
    assert energies[-1] == pytest.approx(ground, abs=1e-2)
