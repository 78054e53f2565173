"""Pauli strings and host-side Pauli sums.

Encoding (src/PauliAlgebra/bitoperations.jl:154, src/PauliAlgebra/utils.jl:2):
qubit q (1-based, as in the reference) occupies bits 2(q-1), 2(q-1)+1 of an
unsigned integer, with I=0, X=1, Y=2, Z=3. A key of W limbs is that integer
split into W little-endian uint64 words.
"""
from __future__ import annotations

import numpy as np

SYMBOLS = {"I": 0, "X": 1, "Y": 2, "Z": 3}
CODES = "IXYZ"
_U64 = (1 << 64) - 1


def key_words(nqubits: int) -> int:
    """Limbs per key: 1 up to 32 qubits, 2 up to 64, 4 up to 128
    (src/PauliAlgebra/bitoperations.jl:8-26 picks UInt64, UInt128, 4-limb NTupleInteger)."""
    if nqubits < 1 or nqubits > 128:
        raise ValueError(f"nqubits must be in 1..128, got {nqubits}")
    return 1 if nqubits <= 32 else (2 if nqubits <= 64 else 4)


def _check_qind(nqubits: int, q: int) -> None:
    if not 1 <= q <= nqubits:
        raise ValueError(f"qubit index {q} is out of range 1..{nqubits}")


def symboltoint(nqubits: int, symbols, qinds) -> int:
    """Integer Pauli string with `symbols` on the 1-based qubits `qinds`
    (src/PauliAlgebra/utils.jl:101-108)."""
    if isinstance(symbols, str) and len(symbols) == 1:
        symbols = [symbols]
    if isinstance(qinds, int):
        qinds = [qinds]
    symbols = list(symbols)
    qinds = list(qinds)
    if len(symbols) != len(qinds):
        raise ValueError("symbols and qinds must have equal length")
    if len(set(qinds)) != len(qinds):
        raise ValueError("qubit indices must be distinct")
    v = 0
    for s, q in zip(symbols, qinds):
        _check_qind(nqubits, q)
        code = SYMBOLS[str(s).upper()]
        v |= code << (2 * (q - 1))
    return v


def int_to_limbs(v: int, W: int) -> list[int]:
    return [(v >> (64 * l)) & _U64 for l in range(W)]


def limbs_to_int(limbs) -> int:
    v = 0
    for l, w in enumerate(limbs):
        v |= int(w) << (64 * l)
    return v


def getpauli(v: int, q: int) -> int:
    """Pauli code on 1-based qubit q."""
    return (v >> (2 * (q - 1))) & 3


def countweight(v: int) -> int:
    w = 0
    while v:
        w += (v & 3) != 0
        v >>= 2
    return w


def containsXorY(v: int) -> bool:
    while v:
        if (v & 3) in (1, 2):
            return True
        v >>= 2
    return False


def pauli_string_text(v: int, nqubits: int) -> str:
    return "".join(CODES[getpauli(v, q)] for q in range(1, nqubits + 1))


class PauliString:
    """A single weighted Pauli string (src/PauliDataTypes/paulistring.jl)."""

    def __init__(self, nqubits: int, symbols, qinds, coeff: float = 1.0):
        self.nqubits = int(nqubits)
        self.term = symboltoint(self.nqubits, symbols, qinds)
        self.coeff = float(coeff)

    def __repr__(self):
        return f"PauliString({self.coeff} * {pauli_string_text(self.term, self.nqubits)})"


class PauliSum:
    """Host Pauli sum: a dict from integer Pauli string to coefficient
    (src/PauliDataTypes/paulisum.jl). Coefficients may be scalars or, for a
    batched sum, length-B arrays."""

    def __init__(self, nqubits: int, terms: dict | None = None):
        self.nqubits = int(nqubits)
        self.terms: dict[int, object] = dict(terms or {})

    @classmethod
    def from_string(cls, pstr: PauliString) -> "PauliSum":
        return cls(pstr.nqubits, {pstr.term: pstr.coeff})

    def add(self, symbols, qinds, coeff: float = 1.0) -> "PauliSum":
        """add!(psum, symbols, qinds, coeff)"""
        t = symboltoint(self.nqubits, symbols, qinds)
        self.terms[t] = self.terms.get(t, 0.0) + coeff
        return self

    def __len__(self):
        return len(self.terms)

    def to_arrays(self, W: int | None = None, batch: int = 1):
        """(keys n x W uint64, coeffs n x batch float64), in dict order."""
        W = W or key_words(self.nqubits)
        n = len(self.terms)
        keys = np.zeros((n, W), dtype=np.uint64)
        coeffs = np.zeros((n, batch), dtype=np.float64)
        for i, (t, c) in enumerate(self.terms.items()):
            keys[i] = int_to_limbs(t, W)
            coeffs[i] = c
        return keys, coeffs

    @classmethod
    def from_arrays(cls, nqubits: int, keys: np.ndarray, coeffs: np.ndarray) -> "PauliSum":
        ps = cls(nqubits)
        coeffs = np.asarray(coeffs)
        for i in range(keys.shape[0]):
            c = coeffs[i]
            ps.terms[limbs_to_int(keys[i])] = float(c[0]) if c.shape == (1,) else c.copy()
        return ps

    def __repr__(self):
        items = list(self.terms.items())[:8]
        body = ", ".join(f"{c} * {pauli_string_text(t, self.nqubits)}" for t, c in items)
        more = "" if len(self.terms) <= 8 else f", ... ({len(self.terms)} terms)"
        return f"PauliSum({body}{more})"
