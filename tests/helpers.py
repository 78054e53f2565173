"""Shared helpers for the tests: random Pauli sums, bitwise comparison."""
from __future__ import annotations

import math

import numpy as np

from pauliprop_hip.paulis import key_words, int_to_limbs


def random_terms(nqubits: int, n: int, rng: np.random.Generator, p_nonid: float = 0.3) -> list[int]:
    """n distinct random Pauli strings on nqubits qubits, each qubit non-identity with prob p_nonid."""
    seen = set()
    out = []
    tries = 0
    while len(out) < n and tries < 20 * n + 100:
        tries += 1
        v = 0
        nz = rng.random(nqubits) < p_nonid
        codes = rng.integers(1, 4, size=nqubits)
        for q in range(nqubits):
            if nz[q]:
                v |= int(codes[q]) << (2 * q)
        if v not in seen:
            seen.add(v)
            out.append(v)
    return out


def terms_to_arrays(terms, nqubits, coeffs):
    W = key_words(nqubits)
    keys = np.array([int_to_limbs(t, W) for t in terms], dtype=np.uint64).reshape(-1, W)
    c = np.asarray(coeffs, dtype=np.float64).reshape(len(terms), -1)
    return keys, c


def mask_of(nqubits, codes: dict) -> int:
    """codes: {qubit (1-based): code}"""
    v = 0
    for q, c in codes.items():
        v |= c << (2 * (q - 1))
    return v


def assert_bitwise(keys, coeffs, rkeys, rcoeffs, what=""):
    assert keys.shape == rkeys.shape, f"{what}: {keys.shape[0]} terms vs oracle {rkeys.shape[0]}"
    assert np.array_equal(keys, rkeys), f"{what}: key sets differ"
    diff = coeffs.view(np.uint64) != rcoeffs.view(np.uint64)
    if diff.any():
        i = int(np.argwhere(diff)[0][0])
        raise AssertionError(f"{what}: {int(diff.sum())} coefficients differ, first at {i}: "
                             f"{coeffs[i]} vs {rcoeffs[i]}")


def rel_close(a, b, tol=1e-12):
    return abs(a - b) <= tol * max(1.0, abs(b))


def fsum_overlap(keys, coeffs) -> np.ndarray:
    """Exactly rounded sum of the coefficients of terms without X or Y."""
    A = np.uint64(0x5555555555555555)
    xy = np.zeros(keys.shape[0], dtype=bool)
    for l in range(keys.shape[1]):
        w = keys[:, l]
        xy |= ((w ^ (w >> np.uint64(1))) & A) != 0
    sel = coeffs[~xy]
    return np.array([math.fsum(sel[:, b]) for b in range(coeffs.shape[1])])
