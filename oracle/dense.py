"""Dense brute force for small qubit counts (numpy): evolves O -> U^dag O U
with explicit 2^n x 2^n matrices and expands the result in the Pauli basis.

TEST INFRASTRUCTURE. It shares no code with the oracle or the GPU path: it
knows only the Pauli matrices and the gate definitions
  PauliRotation(P, theta) = exp(-i theta/2 P)  (src/Gates/paulirotations.jl:11,
                                               heisenbergschrodinger.jl:131-136)
  PauliNoise on qubit q: the Pauli channel whose Heisenberg action damps the
  listed Paulis by (1 - lambda) (src/Gates/noisechannels.jl:24-50).
  CliffordGate: the unitary of each symbol as the reference's own Yao test
  builds it (ext/PauliPropagationYao.jl:8-60, test/test_gates_against_yao.jl:
  117-160): H, S = diag(1, i), X, Y, Z, SX = Rx(pi/2), SY = Ry(pi/2), CNOT and
  CZ with the first qubit as control, SWAP, ZZpihalf = rot(Z (x) Z, -pi/2) =
  exp(+i pi/4 Z (x) Z); Heisenberg action O -> U^dag O U.
  AmplitudeDampingNoise(gamma): the Kraus channel K0 = diag(1, sqrt(1-gamma)),
  K1 = sqrt(gamma) |0><1| (Yao's AmplitudeDampingError, used by the same
  test); Heisenberg action O -> sum_k K_k^dag O K_k.
Qubit q (1-based) is tensor factor q counted from the least significant bit of
the computational-basis index.
"""
from __future__ import annotations

import itertools

import numpy as np

I2 = np.eye(2, dtype=complex)
X = np.array([[0, 1], [1, 0]], dtype=complex)
Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
Z = np.array([[1, 0], [0, -1]], dtype=complex)
PAULI = [I2, X, Y, Z]


def pauli_matrix(codes: dict, n: int) -> np.ndarray:
    """codes: {qubit (1-based): code}. Kronecker order: qubit n first, qubit 1 last."""
    m = np.array([[1.0 + 0j]])
    for q in range(n, 0, -1):
        m = np.kron(m, PAULI[codes.get(q, 0)])
    return m


def term_codes(term: int, n: int) -> dict:
    return {q: (term >> (2 * (q - 1))) & 3 for q in range(1, n + 1) if (term >> (2 * (q - 1))) & 3}


def operator_from_terms(terms: dict, n: int) -> np.ndarray:
    op = np.zeros((2 ** n, 2 ** n), dtype=complex)
    for t, c in terms.items():
        op += c * pauli_matrix(term_codes(t, n), n)
    return op


def rotation_unitary(codes: dict, theta: float, n: int) -> np.ndarray:
    P = pauli_matrix(codes, n)
    return np.cos(theta / 2) * np.eye(2 ** n) - 1j * np.sin(theta / 2) * P


def heisenberg_step_rotation(op, codes, theta, n):
    U = rotation_unitary(codes, theta, n)
    return U.conj().T @ op @ U


def heisenberg_step_noise(op, q, damped_codes, lam, n):
    """Pauli channel acting as P -> (1-lam) P on the damped Paulis of qubit q:
    split op by its Pauli component on qubit q and damp the listed ones."""
    out = np.zeros_like(op)
    for code in range(4):
        comp = _qubit_component(op, q, code, n)
        out += (1 - lam) * comp if code in damped_codes else comp
    return out


def _qubit_component(op, q, code, n):
    """P_code (x) A with A = tr_q(P_code op) / 2: the part of op carrying Pauli
    `code` on qubit q."""
    dim = 2 ** n
    t = op.reshape([2] * (2 * n))
    ax_r, ax_c = n - q, 2 * n - q
    tt = np.moveaxis(t, (ax_r, ax_c), (0, 1))
    P = PAULI[code]
    A = 0.5 * np.einsum("ba,ab...->...", P, tt)
    comp = np.einsum("ab,...->ab...", P, A)
    return np.moveaxis(comp, (0, 1), (ax_r, ax_c)).reshape(dim, dim)


def embed(G, qs, n: int) -> np.ndarray:
    """The 2^n x 2^n matrix of a gate G on the 1-based qubits qs; G's basis
    index holds qubit qs[a] at bit a."""
    dim = 2 ** n
    k = len(qs)
    M = np.zeros((dim, dim), dtype=complex)
    for i in range(dim):
        for j in range(dim):
            rest = all(((i >> (q - 1)) & 1) == ((j >> (q - 1)) & 1) for q in range(1, n + 1) if q not in qs)
            if not rest:
                continue
            si = sum(((i >> (q - 1)) & 1) << a for a, q in enumerate(qs))
            sj = sum(((j >> (q - 1)) & 1) << a for a, q in enumerate(qs))
            M[i, j] = G[si, sj]
    assert k == int(round(np.log2(G.shape[0])))
    return M


def _rot(P, theta):
    return np.cos(theta / 2) * np.eye(P.shape[0]) - 1j * np.sin(theta / 2) * P


def _cnot():
    G = np.zeros((4, 4), dtype=complex)
    for c in range(2):
        for t in range(2):
            G[c | ((t ^ c) << 1), c | (t << 1)] = 1
    return G


def _swap():
    G = np.zeros((4, 4), dtype=complex)
    for a in range(2):
        for b in range(2):
            G[b | (a << 1), a | (b << 1)] = 1
    return G


ZZ = np.kron(Z, Z)
CLIFFORD_UNITARIES = {
    "H": np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2),
    "X": X, "Y": Y, "Z": Z,
    "S": np.diag([1, 1j]).astype(complex),
    "SX": _rot(X, np.pi / 2),
    "SY": _rot(Y, np.pi / 2),
    "CNOT": _cnot(),
    "CZ": np.diag([1, 1, 1, -1]).astype(complex),
    "SWAP": _swap(),
    "ZZPIHALF": _rot(ZZ, -np.pi / 2),
}
CLIFFORD_UNITARIES["SZ"] = CLIFFORD_UNITARIES["S"]


def heisenberg_step_unitary(op, U):
    return U.conj().T @ op @ U


def damping_kraus(gamma: float):
    K0 = np.diag([1.0, np.sqrt(1 - gamma)]).astype(complex)
    K1 = np.array([[0.0, np.sqrt(gamma)], [0.0, 0.0]], dtype=complex)
    return K0, K1


def heisenberg_step_damping(op, q, gamma, n):
    out = np.zeros_like(op)
    for K in damping_kraus(gamma):
        Kf = embed(K, [q], n)
        out += Kf.conj().T @ op @ Kf
    return out


def pauli_expand(op, n, tol=1e-15) -> dict:
    """Coefficients c_t = tr(P_t op) / 2^n for all 4^n Pauli strings above tol."""
    res = {}
    for codes in itertools.product(range(4), repeat=n):
        t = 0
        cd = {}
        for q, c in enumerate(codes, start=1):
            t |= c << (2 * (q - 1))
            if c:
                cd[q] = c
        val = np.trace(pauli_matrix(cd, n) @ op).real / 2 ** n
        if abs(val) > tol:
            res[t] = val
    return res


def zero_state_expectation(op) -> float:
    return float(op[0, 0].real)


def dense_propagate(heisenberg_gates, n, init_terms: dict):
    """heisenberg_gates: list of ("rot", codes, theta), ("noise", q, damped_codes, lam),
    ("clifford", symbol, qs) or ("damp", q, gamma), in propagation (Heisenberg) order."""
    op = operator_from_terms(init_terms, n)
    for g in heisenberg_gates:
        if g[0] == "rot":
            op = heisenberg_step_rotation(op, g[1], g[2], n)
        elif g[0] == "noise":
            op = heisenberg_step_noise(op, g[1], g[2], g[3], n)
        elif g[0] == "clifford":
            op = heisenberg_step_unitary(op, embed(CLIFFORD_UNITARIES[g[1]], list(g[2]), n))
        elif g[0] == "damp":
            op = heisenberg_step_damping(op, g[1], g[2], n)
        else:
            raise ValueError(g[0])
    return op
