"""Gate types of the reference, unchanged in meaning, plus their packing into
the C-ABI gate descriptor `pp_gate` (include/pphip.h).

Reference: src/Gates/paulirotations.jl:11-70 (PauliRotation),
src/Gates/cliffordgates.jl:4-72 (CliffordGate and the default Clifford map),
src/Gates/noisechannels.jl:24-278 (Pauli noise channels, AmplitudeDampingNoise,
the strength check), src/Gates/frozengates.jl:12-45 (FrozenGate, freeze),
src/Gates/heisenbergschrodinger.jl:20-165 (toheisenberg, toschrodinger),
src/Gates/imaginarytime.jl (ImaginaryPauliRotation), src/Gates/miscgates.jl:10-33
(TGate).
"""
from __future__ import annotations

import ctypes
import math

from .paulis import SYMBOLS, symboltoint

# pp_gate kinds (include/pphip.h)
ROTATION, PAULI_NOISE, AMP_DAMPING, CLIFFORD, IMAG_ROTATION = 0, 1, 2, 3, 4


class pp_gate(ctypes.Structure):
    """Mirror of `pp_gate` in include/pphip.h."""

    _fields_ = [
        ("mask", ctypes.c_uint64 * 4),
        ("kind", ctypes.c_int32),
        ("param", ctypes.c_int32),
        ("value", ctypes.c_double),
        ("qubit", ctypes.c_int32),
        ("damp_codes", ctypes.c_uint32),
        ("nq", ctypes.c_int32),
        ("qubits", ctypes.c_int32 * 2),
        ("lut", ctypes.c_int32),
        ("flags", ctypes.c_int32),
    ]


class Gate:
    parametrized = False


class PauliRotation(Gate):
    """exp(-i theta/2 P) for the Pauli string P = `symbols` on `qinds` (1-based)."""

    parametrized = True

    def __init__(self, symbols, qinds):
        if isinstance(symbols, str) and len(symbols) == 1:
            symbols = [symbols]
        if isinstance(qinds, int):
            qinds = [qinds]
        self.symbols = [str(s).upper() for s in symbols]
        self.qinds = [int(q) for q in qinds]
        if len(self.symbols) != len(self.qinds):
            raise ValueError("symbols and qinds must have equal length")
        for s in self.symbols:
            if s not in SYMBOLS:
                raise ValueError(f"unknown Pauli symbol {s}")

    def mask(self, nqubits: int) -> int:
        return symboltoint(nqubits, self.symbols, self.qinds)

    def __repr__(self):
        return f"PauliRotation({self.symbols}, {self.qinds})"


class ImaginaryPauliRotation(PauliRotation):
    """exp(-tau/2 P) . exp(-tau/2 P) on states (src/Gates/imaginarytime.jl):
    terms commuting with P branch into cosh(tau) and sinh(tau) parts. Defined
    in the Schrodinger picture only (heisenbergschrodinger.jl:75-79)."""

    def __repr__(self):
        return f"ImaginaryPauliRotation({self.symbols}, {self.qinds})"


# The default Clifford map: for each gathered Pauli index (qubit 1 in the low
# two bits) the new partial Pauli string and the sign
# (src/Gates/cliffordgates.jl:47-59). The table is data the gate semantics
# fix, so it is kept entry for entry.
CLIFFORD_MAP = {
    "H": [(0x0, 1), (0x3, 1), (0x2, -1), (0x1, 1)],
    "X": [(0x0, 1), (0x1, 1), (0x2, -1), (0x3, -1)],
    "Y": [(0x0, 1), (0x1, -1), (0x2, 1), (0x3, -1)],
    "Z": [(0x0, 1), (0x1, -1), (0x2, -1), (0x3, 1)],
    "SX": [(0x0, 1), (0x1, 1), (0x3, -1), (0x2, 1)],
    "SY": [(0x0, 1), (0x3, 1), (0x2, 1), (0x1, -1)],
    "S": [(0x0, 1), (0x2, -1), (0x1, 1), (0x3, 1)],
    "CNOT": [(0x0, 1), (0x5, 1), (0x6, 1), (0x3, 1), (0x4, 1), (0x1, 1), (0x2, 1), (0x7, 1),
             (0xB, 1), (0xE, 1), (0xD, -1), (0x8, 1), (0xF, 1), (0xA, -1), (0x9, 1), (0xC, 1)],
    "CZ": [(0x0, 1), (0xD, 1), (0xE, 1), (0x3, 1), (0x7, 1), (0xA, 1), (0x9, -1), (0x4, 1),
           (0xB, 1), (0x6, -1), (0x5, 1), (0x8, 1), (0xC, 1), (0x1, 1), (0x2, 1), (0xF, 1)],
    "SWAP": [(0x0, 1), (0x4, 1), (0x8, 1), (0xC, 1), (0x1, 1), (0x5, 1), (0x9, 1), (0xD, 1),
             (0x2, 1), (0x6, 1), (0xA, 1), (0xE, 1), (0x3, 1), (0x7, 1), (0xB, 1), (0xF, 1)],
    "ZZPIHALF": [(0x0, 1), (0xE, 1), (0xD, -1), (0x3, 1), (0xB, 1), (0x5, 1), (0x6, 1), (0x8, 1),
                 (0x7, -1), (0x9, 1), (0xA, 1), (0x4, -1), (0xC, 1), (0x2, 1), (0x1, -1), (0xF, 1)],
}
CLIFFORD_MAP["SZ"] = CLIFFORD_MAP["S"]


def transposecliffordmap(table):
    """The transposed map: entry q is (p, s) where the map sends p to (q, s)
    (src/Gates/cliffordgates.jl:94-98)."""
    out = [None] * len(table)
    for p, (q, sgn) in enumerate(table):
        out[q] = (p, sgn)
    return out


class CliffordGate(Gate):
    def __init__(self, symbol: str, qinds):
        if isinstance(qinds, int):
            qinds = [qinds]
        self.symbol = str(symbol).upper()
        self.qinds = [int(q) for q in qinds]
        if self.symbol not in CLIFFORD_MAP:
            raise ValueError(f"unknown Clifford gate {symbol}")
        table = CLIFFORD_MAP[self.symbol]
        if len(table) != 4 ** len(self.qinds):
            need = 1 if len(table) == 4 else 2
            raise ValueError(f"Clifford gate {symbol} acts on {need} qubit(s), got {len(self.qinds)}")

    def __repr__(self):
        return f"CliffordGate({self.symbol}, {self.qinds})"


def _check_noise_strength(name: str, lam: float) -> None:
    if not (0.0 <= lam <= 1.0):
        raise ValueError(f"{name} parameter must be between 0 and 1. Got {lam}.")


class PauliNoise(Gate):
    """Damps the Pauli codes in `damped` on one qubit by 1 - lambda
    (src/Propagation/specializations.jl:218-238)."""

    parametrized = True
    damped: tuple = ()

    def __init__(self, qind: int):
        self.qind = int(qind)

    @property
    def damp_codes(self) -> int:
        return sum(1 << c for c in self.damped)

    def __repr__(self):
        return f"{type(self).__name__}({self.qind})"


def _noise_type(name, damped):
    cls = type(name, (PauliNoise,), {"damped": damped})

    def make(qind: int, lam: float | None = None):
        g = cls(qind)
        if lam is None:
            return g
        _check_noise_strength(name, lam)
        return FrozenGate(g, lam)

    make.__name__ = name
    make.gate_type = cls
    return make


# isdamped tables (src/Gates/noisechannels.jl:48-243)
DepolarizingNoise = _noise_type("DepolarizingNoise", (1, 2, 3))
PauliXNoise = _noise_type("PauliXNoise", (2, 3))
PauliYNoise = _noise_type("PauliYNoise", (1, 3))
PauliZNoise = _noise_type("PauliZNoise", (1, 2))
DephasingNoise = PauliZNoise
PauliXDamping = _noise_type("PauliXDamping", (1,))
PauliYDamping = _noise_type("PauliYDamping", (2,))
PauliZDamping = _noise_type("PauliZDamping", (3,))


class AmplitudeDampingNoise(Gate):
    parametrized = True

    def __init__(self, qind: int):
        self.qind = int(qind)

    def __repr__(self):
        return f"AmplitudeDampingNoise({self.qind})"


class TGate(Gate):
    """The T gate on qubit `qind`: a PauliRotation(:Z, qind) with angle pi/4
    (src/Gates/miscgates.jl:10-33, src/Propagation/specializations.jl:293-308)."""

    def __init__(self, qind: int):
        self.qind = int(qind)
        if self.qind < 1:
            raise ValueError("qubit index must be >= 1")

    def __repr__(self):
        return f"TGate({self.qind})"


class FrozenGate(Gate):
    """A parametrized gate with its parameter fixed (src/Gates/frozengates.jl:12)."""

    parametrized = False

    def __init__(self, gate: Gate, parameter: float):
        if not gate.parametrized:
            raise ValueError("only parametrized gates can be frozen")
        self.gate = gate
        self.parameter = float(parameter)

    def __repr__(self):
        return f"FrozenGate({self.gate!r}, {self.parameter})"


def freeze(gate: Gate, parameter: float) -> FrozenGate:
    return FrozenGate(gate, parameter)


def countparameters(circuit) -> int:
    return sum(1 for g in circuit if g.parametrized)


def _inner(gate):
    return gate.gate if isinstance(gate, FrozenGate) else gate


def toheisenberg(circuit, thetas):
    """Reverse gates and parameters; angles are unchanged
    (src/Gates/heisenbergschrodinger.jl:20-49). ImaginaryPauliRotation has no
    Heisenberg form (:75-79)."""
    n = countparameters(circuit)
    if thetas is not None and len(thetas) != n:
        raise ValueError(f"the circuit has {n} parameters, got {len(thetas)}")
    for g in circuit:
        if isinstance(_inner(g), ImaginaryPauliRotation):
            raise RuntimeError("ImaginaryPauliRotation gates are currently not defined in the Heisenberg picture.")
    return list(reversed(circuit)), (None if thetas is None else list(reversed(list(thetas))))


def toschrodinger(circuit, thetas):
    """Schrodinger-picture gates in circuit order (src/Gates/heisenbergschrodinger.jl:87-165):
    a PauliRotation's angle changes sign, a Clifford takes its transposed map,
    Pauli noise, amplitude damping and ImaginaryPauliRotation are unchanged;
    a TGate has no Schrodinger form here, as in the reference."""
    n = countparameters(circuit)
    if thetas is not None and len(thetas) != n:
        raise ValueError(f"the circuit has {n} parameters, got {len(thetas)}")
    out, th, k = [], [], 0
    for g in circuit:
        inner = _inner(g)
        if isinstance(inner, TGate):
            raise NotImplementedError("toschrodinger is not implemented for TGate")
        if isinstance(inner, CliffordGate):
            sym = inner.symbol + "_TRANSPOSE"
            if sym not in CLIFFORD_MAP:
                CLIFFORD_MAP[sym] = transposecliffordmap(CLIFFORD_MAP[inner.symbol])
            out.append(CliffordGate(sym, inner.qinds))
            continue
        flip = isinstance(inner, PauliRotation) and not isinstance(inner, ImaginaryPauliRotation)
        if isinstance(g, FrozenGate):
            out.append(FrozenGate(inner, -g.parameter if flip else g.parameter))
        else:
            out.append(g)
            if g.parametrized:
                if thetas is not None:
                    v = thetas[k]
                    th.append(-v if flip else v)
                k += 1
    return out, (None if thetas is None else th)


def _qubit0(nqubits: int, q: int) -> int:
    if not 1 <= q <= nqubits:
        raise ValueError(f"qubit index {q} is out of range 1..{nqubits}")
    return q - 1


def clifford_table(symbol: str, qubits0, W: int):
    """The 16-entry table of a Clifford on 0-based qubits: per gathered index
    idx = code(q0) | code(q1) << 2, the XOR change of the whole key (W limbs,
    limb 0 first) and the sign (entries past 4^nq are identities)."""
    table = CLIFFORD_MAP[str(symbol).upper()]
    changes: list[int] = []
    signs: list[float] = []
    for idx in range(16):
        if idx < len(table):
            new, sign = table[idx]
            change = 0
            for a, q in enumerate(qubits0):
                old_p = (idx >> (2 * a)) & 3
                new_p = (new >> (2 * a)) & 3
                change |= (old_p ^ new_p) << (2 * q)
            changes.extend((change >> (64 * l)) & ((1 << 64) - 1) for l in range(W))
            signs.append(float(sign))
        else:
            changes.extend([0] * W)
            signs.append(1.0)
    return changes, signs


def pack_circuit(circuit, nqubits: int, W: int, param_offset=None, normalize_coeffs: bool = True):
    """Pack a gate list in propagation order (Heisenberg or Schrodinger) into
    pp_gate descriptors; normalize_coeffs sets the flag of every
    ImaginaryPauliRotation.

    Parametrized gates read parameter rows in order (row k is the k-th
    parametrized gate of the list). Returns (gates array, Clifford table
    changes as flat uint64 list, signs list, number of tables).
    """
    arr = (pp_gate * max(len(circuit), 1))()
    lut_changes: list[int] = []
    lut_signs: list[float] = []
    lut_slots: dict[tuple, int] = {}
    k = 0
    for gi, gate in enumerate(circuit):
        d = arr[gi]
        d.param = -1
        inner, frozen = (gate.gate, gate.parameter) if isinstance(gate, FrozenGate) else (gate, None)
        if isinstance(inner, TGate):
            inner, frozen = PauliRotation("Z", inner.qind), math.pi / 4
        if frozen is None and inner.parametrized:
            d.param = k
            k += 1
        else:
            d.value = 0.0 if frozen is None else frozen
        if isinstance(inner, PauliRotation):
            d.kind = IMAG_ROTATION if isinstance(inner, ImaginaryPauliRotation) else ROTATION
            d.flags = 1 if (d.kind == IMAG_ROTATION and normalize_coeffs) else 0
            m = inner.mask(nqubits)
            for l in range(W):
                d.mask[l] = (m >> (64 * l)) & ((1 << 64) - 1)
        elif isinstance(inner, PauliNoise):
            d.kind = PAULI_NOISE
            d.qubit = _qubit0(nqubits, inner.qind)
            d.damp_codes = inner.damp_codes
            if frozen is not None:
                _check_noise_strength(type(inner).__name__, frozen)
        elif isinstance(inner, AmplitudeDampingNoise):
            d.kind = AMP_DAMPING
            d.qubit = _qubit0(nqubits, inner.qind)
            if frozen is not None:
                _check_noise_strength("AmplitudeDampingNoise", frozen)
        elif isinstance(inner, CliffordGate):
            d.kind = CLIFFORD
            d.nq = len(inner.qinds)
            qs = [_qubit0(nqubits, q) for q in inner.qinds]
            for a, q in enumerate(qs):
                d.qubits[a] = q
            slot_key = (inner.symbol, tuple(qs))
            if slot_key not in lut_slots:
                lut_slots[slot_key] = len(lut_slots)
                ch, sg = clifford_table(inner.symbol, qs, W)
                lut_changes.extend(ch)
                lut_signs.extend(sg)
            d.lut = lut_slots[slot_key]
        else:
            raise NotImplementedError(f"gate {gate!r} is not supported by this backend")
    return arr, lut_changes, lut_signs, len(lut_slots)
