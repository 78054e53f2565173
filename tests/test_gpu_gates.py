"""GPU parity, single gates: libpphip.so through the C ABI against the CPU
oracle on identical inputs, bit for bit (keys and FP64 coefficients)."""
import math

import numpy as np
import pytest

from helpers import assert_bitwise, mask_of, random_terms, terms_to_arrays, fsum_overlap
from oracle import OracleSum
from pauliprop_hip import HIPPauliSum
from pauliprop_hip.paulis import key_words, int_to_limbs

pytestmark = pytest.mark.gpu

# (nqubits, {qubit: code}) -- low, middle, top, limb straddle, wrap-around, multi-run masks
MASKS = [
    (32, {1: 1}), (32, {16: 3}), (32, {32: 2}), (32, {1: 3, 2: 3}), (32, {32: 3, 1: 3}),
    (32, {5: 1, 9: 1}), (32, {31: 2, 32: 2}), (32, {1: 1, 10: 2, 20: 3}),
    (64, {33: 1}), (64, {32: 3, 33: 3}), (64, {1: 2, 64: 2}), (64, {17: 1, 40: 2, 63: 3}),
    (127, {127: 1}), (127, {63: 3, 65: 3}), (127, {64: 3}), (127, {1: 3, 127: 3}), (127, {30: 1, 95: 2}),
    (127, {2: 1, 4: 2, 6: 3, 8: 1, 70: 2, 100: 3, 126: 1}),
]


def _pair(nq, n, seed, p=0.3, batch=1):
    rng = np.random.default_rng(seed)
    terms = random_terms(nq, n, rng, p)
    coeffs = rng.standard_normal((len(terms), batch))
    keys, c = terms_to_arrays(terms, nq, coeffs)
    dev = HIPPauliSum(nq, batch=batch, capacity=max(1024, len(terms)))
    dev.upload(keys, c)
    ref = OracleSum(key_words(nq), batch).set(keys, c)
    return dev, ref


@pytest.mark.parametrize("nq,codes", MASKS)
@pytest.mark.parametrize("n", [1, 5, 3000, 120000])
def test_rotation_weight_only(nq, codes, n):
    print(f"rotation nq={nq} codes={codes} n={n}", flush=True)
    dev, ref = _pair(nq, n, seed=n + nq)
    m = mask_of(nq, codes)
    theta = 0.37
    W = key_words(nq)
    for step in range(3):
        dev.apply_rotation(m, theta, max_weight=12, min_abs_coeff=0.0)
        ref.rotation(int_to_limbs(m, W), math.cos(theta), math.sin(theta), max_weight=12, min_abs=0.0)
        k, c = dev.download()
        rk, rc = ref.get()
        assert_bitwise(k, c, rk, rc, f"step {step}")


@pytest.mark.parametrize("nq,codes", [(32, {16: 1}), (64, {20: 3, 21: 3}), (127, {63: 1, 64: 2})])
@pytest.mark.parametrize("min_abs", [1e-3, 0.3])
def test_rotation_coefficient_truncation(nq, codes, min_abs):
    print(f"truncation nq={nq} codes={codes} min_abs={min_abs}", flush=True)
    dev, ref = _pair(nq, 60000, seed=7)
    m = mask_of(nq, codes)
    W = key_words(nq)
    for theta in (0.2, 1.1, -0.4):
        dev.apply_rotation(m, theta, max_weight=math.inf, min_abs_coeff=min_abs)
        ref.rotation(int_to_limbs(m, W), math.cos(theta), math.sin(theta), min_abs=min_abs)
        k, c = dev.download()
        rk, rc = ref.get()
        assert_bitwise(k, c, rk, rc, f"theta {theta}")


@pytest.mark.parametrize("nq", [32, 64, 127])
def test_noise_and_overlap(nq):
    print(f"noise nq={nq}", flush=True)
    dev, ref = _pair(nq, 50000, seed=3, p=0.2)
    for q, codes, lam in [(1, 0b1110, 0.1), (nq, 0b0110, 0.5), (nq // 2, 0b1000, 0.9)]:
        dev.apply_pauli_noise(q, codes, lam, max_weight=10, min_abs_coeff=0.05)
        ref.pauli_noise(q - 1, codes, 1 - lam, max_weight=10, min_abs=0.05)
        k, c = dev.download()
        rk, rc = ref.get()
        assert_bitwise(k, c, rk, rc, f"noise q={q}")
    ov = dev.overlapwithzero()
    exact = fsum_overlap(*dev.download())[0]
    assert abs(ov - exact) <= 1e-14 * max(1.0, abs(exact))


def test_untouched_gate_leaves_sum_alone():
    # a gate touching no term neither merges nor truncates (src/Base/Merge/xortailmerge.jl:121-123)
    nq = 8
    dev = HIPPauliSum(nq)
    keys, c = terms_to_arrays([mask_of(nq, {1: 3}), mask_of(nq, {2: 3, 3: 3})], nq, [1e-9, 0.5])
    dev.upload(keys, c)
    st = dev.apply_rotation(mask_of(nq, {8: 1}), 0.3, min_abs_coeff=1e-3)
    assert st.n_touched == 0 and len(dev) == 2
    st = dev.apply_rotation(mask_of(nq, {1: 1}), 0.3, min_abs_coeff=1e-3)
    assert st.n_touched == 1
    k, cc = dev.download()
    # the 1e-9 term (kept through the untouched gate) is truncated with its branch now
    assert len(k) == 1 and int(k[0, 0]) == mask_of(nq, {2: 3, 3: 3})


def test_empty_sum_and_errors():
    dev = HIPPauliSum(40)
    assert len(dev) == 0
    dev.apply_rotation(mask_of(40, {3: 1}), 0.1)
    assert dev.overlapwithzero() == 0.0
    with pytest.raises(ValueError):
        dev.apply_rotation(1 << 90, 0.1)  # qubit 46 of a 40-qubit sum
    keys, c = terms_to_arrays([5, 5], 40, [1.0, 2.0])
    with pytest.raises(ValueError):
        dev.upload(keys, c)


@pytest.mark.parametrize("mode", ["tiny_spin_limit", "forced_rerun", "dynamic_tiles"])
def test_tile_claim_modes_are_bit_identical(mode):
    # static tiles whose waits run out (a one-poll limit, where timing decides;
    # or reported as run out by every static-tile launch) stop the gate, which is
    # run again with atomic tile claims; every mode must match the oracle
    print(f"tile mode {mode}", flush=True)
    nq = 64
    dev, ref = _pair(nq, 200000, seed=11)
    if mode == "tiny_spin_limit":
        dev.tune(spin_limit=1)
    elif mode == "forced_rerun":
        dev.tune(force_rerun=True)
    else:
        dev.tune(dynamic_tiles=True)
    W = key_words(nq)
    for codes in ({20: 1}, {20: 3, 21: 3}, {5: 2, 60: 1}, {64: 3}):
        m = mask_of(nq, codes)
        dev.apply_rotation(m, 0.7, max_weight=10, min_abs_coeff=1e-4)
        ref.rotation(int_to_limbs(m, W), math.cos(0.7), math.sin(0.7), max_weight=10, min_abs=1e-4)
        assert_bitwise(*dev.download(), *ref.get(), f"{mode} {codes}")
    dev.apply_pauli_noise(7, 0b1110, 0.2, max_weight=10, min_abs_coeff=1e-3)
    ref.pauli_noise(6, 0b1110, 0.8, max_weight=10, min_abs=1e-3)
    assert_bitwise(*dev.download(), *ref.get(), f"{mode} noise")
    fb = dev.profile()["fallbacks"]
    print("fallbacks", fb, flush=True)
    if mode == "forced_rerun":
        assert fb > 0, "no gate was sent to the atomic-claim rerun"


@pytest.mark.parametrize("nq,qubit", [(32, 1), (32, 17), (64, 33), (127, 64), (127, 127)])
@pytest.mark.parametrize("n", [1, 5, 3000, 120000])
@pytest.mark.parametrize("min_abs", [0.0, 1e-2])
def test_amplitude_damping(nq, qubit, n, min_abs):
    # Z branches into I (gamma c) and keeps (1 - gamma) c; X, Y scale by sqrt(1 - gamma)
    # (src/Propagation/specializations.jl:258-276)
    print(f"damping nq={nq} qubit={qubit} n={n} min_abs={min_abs}", flush=True)
    dev, ref = _pair(nq, n, seed=3 * n + qubit)
    for gamma in (0.1, 0.37):
        dev.apply_amplitude_damping(qubit, gamma, max_weight=10, min_abs_coeff=min_abs)
        ref.amplitude_damping(qubit - 1, gamma, max_weight=10, min_abs=min_abs)
        assert_bitwise(*dev.download(), *ref.get(), f"gamma {gamma}")


def test_amplitude_damping_without_z_and_without_support():
    # only X/Y on the qubit: coefficients change and everything is truncated, nothing
    # is created; no support at all: the sum is left as it is (no truncation)
    nq = 40
    keys, c = terms_to_arrays([1 << 2 * 4, 2 << 2 * 4, (1 << 2 * 4) | (3 << 2 * 7)], nq, [0.5, 1e-3, 0.2])
    for qubit in (5, 20):
        dev = HIPPauliSum(nq)
        dev.upload(keys, c)
        ref = OracleSum(key_words(nq), 1).set(keys, c)
        dev.apply_amplitude_damping(qubit, 0.5, min_abs_coeff=1e-2)
        ref.amplitude_damping(qubit - 1, 0.5, min_abs=1e-2)
        assert_bitwise(*dev.download(), *ref.get(), f"qubit {qubit}")
    assert len(dev.download()[0]) == 3  # qubit 20 carries no Pauli: untouched, nothing truncated


def test_amplitude_damping_arguments():
    dev = HIPPauliSum(40)
    with pytest.raises(ValueError):
        dev.apply_amplitude_damping(41, 0.1)
    with pytest.raises(ValueError):
        dev.apply_amplitude_damping(3, 1.5)


@pytest.mark.parametrize("nq,symbol,qinds", [
    (32, "H", [1]), (32, "S", [17]), (32, "SX", [32]), (64, "H", [33]), (127, "SY", [127]),
    (32, "CNOT", [3, 4]), (32, "CZ", [20, 9]), (64, "SWAP", [31, 34]), (127, "CNOT", [64, 63]),
    (64, "CNOT", [1, 40]), (127, "ZZPIHALF", [2, 120]),  # windows wider than 32 bits: limb rounds
])
@pytest.mark.parametrize("n", [1, 5, 3000, 120000])
@pytest.mark.parametrize("max_weight", [math.inf, 9])
def test_clifford(nq, symbol, qinds, n, max_weight):
    # map every term, truncate only under a weight limit, and sort again on the device;
    # the oracle leaves its sum unsorted (src/Base/Primitives/map.jl:63), so it is compared sorted
    from pauliprop_hip.gates import clifford_table
    print(f"clifford {symbol}{qinds} nq={nq} n={n} max_weight={max_weight}", flush=True)
    dev, ref = _pair(nq, n, seed=n + 7 * nq)
    ch, sg = clifford_table(symbol, [q - 1 for q in qinds], key_words(nq))
    for rep in range(2):
        dev.apply_clifford(symbol, qinds, max_weight=max_weight, min_abs_coeff=1e-3)
        ref.clifford([q - 1 for q in qinds], ch, sg, max_weight=max_weight, min_abs=1e-3)
        assert_bitwise(*dev.download(), *ref.get_sorted(), f"{symbol} application {rep}")


def test_clifford_batched_and_arguments():
    from pauliprop_hip.gates import clifford_table
    nq = 64
    dev, ref = _pair(nq, 50000, seed=5, batch=4)
    ch, sg = clifford_table("CNOT", [9, 10], key_words(nq))
    dev.apply_clifford("CNOT", [10, 11])
    ref.clifford([9, 10], ch, sg)
    assert_bitwise(*dev.download(), *ref.get_sorted(), "batched CNOT")
    with pytest.raises(ValueError):
        dev.apply_clifford("CNOT", [3, 3])
    with pytest.raises(ValueError):
        dev.apply_clifford("H", [65])


def test_load_merge_known_answer():
    # test/test_mergeandtruncate.jl:8-20: terms [1,1,2,2,3], coefficients
    # [.60, -.55, .06, .06, 1.0], min_abs_coeff 0.1 -> {2: 0.12, 3: 1.0}
    dev = HIPPauliSum(4)
    keys = np.array([[1], [1], [2], [2], [3]], dtype=np.uint64)
    dev.load_merge(keys, np.array([0.60, -0.55, 0.06, 0.06, 1.0]), min_abs_coeff=0.1)
    k, c = dev.download()
    assert k[:, 0].tolist() == [2, 3]
    assert c[:, 0].tolist() == [0.06 + 0.06, 1.0]


@pytest.mark.parametrize("nq", [20, 64, 127])
@pytest.mark.parametrize("n", [1, 7, 5000, 100000])
@pytest.mark.parametrize("B", [1, 3])
def test_load_merge_with_repeats(nq, n, B):
    # strings drawn from a small pool so that runs of equal strings occur, in random order
    print(f"load_merge nq={nq} n={n} B={B}", flush=True)
    rng = np.random.default_rng(n + nq + B)
    pool = random_terms(nq, max(1, n // 4), rng, 0.3)
    W = key_words(nq)
    picks = rng.integers(0, len(pool), size=n)
    keys = np.array([int_to_limbs(pool[i], W) for i in picks], dtype=np.uint64).reshape(-1, W)
    coeffs = rng.standard_normal((n, B))
    for mw, mac in ((math.inf, 0.0), (8, 0.05)):
        dev = HIPPauliSum(nq, batch=B)
        dev.load_merge(keys, coeffs, max_weight=mw, min_abs_coeff=mac)
        ref = OracleSum(W, B).load_merge(keys, coeffs, max_weight=mw, min_abs=mac)
        assert_bitwise(*dev.download(), *ref.get(), f"max_weight {mw} min_abs {mac}")


def test_clone_is_an_independent_copy():
    # pp_clone: what deepcopy of the sum gives (src/Base/propagate.jl:3); the
    # copy and the original evolve separately, and the copy outlives the original
    import copy
    print("pp_clone", flush=True)
    nq = 64
    dev, ref = _pair(nq, 50000, seed=3)
    W = key_words(nq)
    c = copy.deepcopy(dev)
    assert_bitwise(*c.download(), *dev.download(), "clone")
    m = mask_of(nq, {10: 1, 11: 2})
    c.apply_rotation(m, 0.4, max_weight=9, min_abs_coeff=1e-5)
    assert_bitwise(*dev.download(), *ref.get(), "original after the copy changed")
    ref.rotation(int_to_limbs(m, W), math.cos(0.4), math.sin(0.4), max_weight=9, min_abs=1e-5)
    dev.close()
    assert_bitwise(*c.download(), *ref.get(), "copy after a gate")
    c2 = c.clone()
    c.close()
    assert len(c2) == len(ref)
