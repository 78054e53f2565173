"""pp_overlap_noalloc (the overlap without device allocation) through the C
ABI, against pp_overlap on the same handles: every overlap kind, two calls in
a row with the second on a smaller sum (the shared static scratch keeps the
first call's partials), and the column limit of its scratch."""
import ctypes

import numpy as np
import pytest

import pauliprop_hip as pp
from pauliprop_hip import _lib

pytestmark = pytest.mark.gpu

_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)


def _sum(nq, n, B, seed):
    """n distinct strings on nq <= 64 qubits: random ones, ones of I and Z only
    (they count for |0..0> and basis states), of I and X only (they count for
    |+..+>), and the identity (the maximally mixed state)."""
    rng = np.random.default_rng(seed)
    W = pp.key_words(nq)
    m = int(n * 1.1) + 8
    codes = rng.integers(0, 4, size=(m, nq)).astype(np.uint64)
    fam = np.arange(m) % 3
    codes[fam == 1] = np.where(codes[fam == 1] & np.uint64(1), np.uint64(3), np.uint64(0))
    codes[fam == 2] &= np.uint64(1)
    keys = np.zeros((m + 1, W), dtype=np.uint64)  # the last row stays the identity
    for q in range(nq):
        keys[:m, q // 32] |= codes[:, q] << np.uint64(2 * (q % 32))
    keys = np.unique(keys, axis=0)[:n]  # ascending: the identity comes first
    coeffs = rng.uniform(-1, 1, size=(keys.shape[0], B))
    h = pp.HIPPauliSum(nq, batch=B, capacity=max(1024, keys.shape[0]))
    h.upload(keys, coeffs)
    return h


def _overlaps(h, fn, kind, ones_int):
    ones = np.ascontiguousarray(pp.paulis.int_to_limbs(ones_int, h.W), dtype=np.uint64)
    out = np.zeros(h.batch)
    _lib.check(fn(h._h, kind, ones.ctypes.data_as(_U64P), out.ctypes.data_as(_DP)))
    return out


@pytest.mark.parametrize("B", [1, 3, 64])
def test_noalloc_overlap_equals_overlap_two_calls(B):
    L = _lib.lib()
    print(f"noalloc overlap B={B}", flush=True)
    big = _sum(40, 300000 if B == 1 else 20000, B, seed=B)
    small = _sum(40, 1500, B, seed=100 + B)
    ones_int = sum(1 << (2 * q) for q in (0, 3, 17, 39))
    for h in (big, small):  # the second call runs on a smaller sum
        for kind in (_lib.PP_OVERLAP_ZERO, _lib.PP_OVERLAP_PLUS, _lib.PP_OVERLAP_COMPUTATIONAL,
                     _lib.PP_OVERLAP_MAXMIXED):
            want = _overlaps(h, L.pp_overlap, kind, ones_int)
            got = _overlaps(h, L.pp_overlap_noalloc, kind, ones_int)
            np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * max(1.0, float(np.abs(want).max())))


def test_noalloc_overlap_after_empty_and_column_limit():
    L = _lib.lib()
    h = pp.HIPPauliSum(10, batch=2)
    assert np.array_equal(_overlaps(h, L.pp_overlap_noalloc, _lib.PP_OVERLAP_ZERO, 0), np.zeros(2))
    wide = pp.HIPPauliSum(10, batch=65)
    with pytest.raises(ValueError):
        _overlaps(wide, L.pp_overlap_noalloc, _lib.PP_OVERLAP_ZERO, 0)
