"""Bit algebra of the oracle against a brute-force per-qubit truth that shares
no code with it: 2x2 Pauli matrix products (as test/test_wideintegers.jl:12-22
checks the limb operations against a per-qubit table)."""
import numpy as np
import pytest

import oracle as orc
from dense import PAULI
from pauliprop_hip.paulis import int_to_limbs


def _product_phase(a: int, b: int):
    """P_a P_b = i^e P_(a xor b) for single-qubit codes: returns e (mod 4)."""
    prod = PAULI[a] @ PAULI[b]
    c = a ^ b
    ref = PAULI[c]
    # find phase: prod = ph * ref
    k = np.flatnonzero(np.abs(ref.ravel()) > 0)[0]
    ph = prod.ravel()[k] / ref.ravel()[k]
    for e in range(4):
        if abs(ph - 1j ** e) < 1e-12:
            return e
    raise AssertionError("no phase")


TABLE = [[_product_phase(a, b) for b in range(4)] for a in range(4)]


def test_single_qubit_table_is_the_pauli_group():
    # I=0 X=1 Y=2 Z=3: XY = iZ, YX = -iZ, ...
    assert TABLE[1][2] == 1 and TABLE[2][1] == 3 and TABLE[2][3] == 1 and TABLE[3][1] == 1


def _truth(p1: int, p2: int, nq: int):
    e = 0
    anti = 0
    for q in range(nq):
        a = (p1 >> (2 * q)) & 3
        b = (p2 >> (2 * q)) & 3
        e += TABLE[a][b]
        anti ^= int(a != 0 and b != 0 and a != b)
    return anti, e & 3


def _call(fn, a, b, W):
    A = np.array(int_to_limbs(a, W), dtype=np.uint64)
    Bv = np.array(int_to_limbs(b, W), dtype=np.uint64)
    return fn(orc.u64p(A), orc.u64p(Bv), W)


def test_exhaustive_two_qubit_pairs():
    L = orc.lib()
    for p1 in range(16):
        for p2 in range(16):
            anti, e = _truth(p1, p2, 2)
            assert _call(L.ppo_anticommutes, p1, p2, 1) == anti
            assert _call(L.ppo_sign_exponent, p1, p2, 1) == e


@pytest.mark.parametrize("nq,W", [(32, 1), (64, 2), (65, 4), (127, 4), (128, 4)])
def test_random_wide_strings(nq, W):
    # random strings over 65..128 qubits, including masks straddling limb boundaries
    # (test/test_wideintegers.jl:197-301)
    L = orc.lib()
    rng = np.random.default_rng(nq)
    for _ in range(400):
        codes1 = rng.integers(0, 4, nq) * (rng.random(nq) < 0.5)
        codes2 = rng.integers(0, 4, nq) * (rng.random(nq) < 0.5)
        p1 = sum(int(c) << (2 * q) for q, c in enumerate(codes1))
        p2 = sum(int(c) << (2 * q) for q, c in enumerate(codes2))
        anti, e = _truth(p1, p2, nq)
        assert _call(L.ppo_anticommutes, p1, p2, W) == anti
        assert _call(L.ppo_sign_exponent, p1, p2, W) == e
        w = int(np.count_nonzero(codes2))
        A = np.array(int_to_limbs(p2, W), dtype=np.uint64)
        assert L.ppo_weight(orc.u64p(A), W) == w
        assert L.ppo_contains_xy(orc.u64p(A), W) == int(any(c in (1, 2) for c in codes2))


def test_rotation_sign_rule_closed_forms():
    # test/test_paulirotations.jl:4-45: RX on Z -> cos Z + sin Y; RX on Y -> cos Y - sin Z;
    # RZ on X -> cos X - sin Y; RYY on Z2 -> cos IZ - sin YX
    th = 0.3
    c, s = np.cos(th), np.sin(th)

    def run(mask, term, W=1):
        o = orc.OracleSum(W, 1).set(np.array([int_to_limbs(term, W)], dtype=np.uint64), np.array([1.0]))
        o.rotation(int_to_limbs(mask, W), c, s, min_abs=0.0)
        k, v = o.get()
        return {int(a[0]): float(b[0]) for a, b in zip(k, v)}

    X, Y, Z = 1, 2, 3
    assert run(X, Z) == {Z: c, Y: s}
    assert run(X, Y) == {Y: c, Z: -s}
    assert run(Z, X) == {X: c, Y: -s}
    YY = Y | (Y << 2)
    IZ = Z << 2
    YX = Y | (X << 2)
    assert run(YY, IZ) == {IZ: c, YX: -s}
