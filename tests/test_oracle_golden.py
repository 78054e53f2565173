"""The oracle against the reference's published values and known answers.

These pin the CPU restatement (oracle/ppo.c); the GPU is then held to the
oracle bit for bit."""
import math

import numpy as np
import pytest

from helpers import fsum_overlap
from oracle import OracleSum, oracle_propagate
import oracle as orc
import pauliprop_hip as pp
from pauliprop_hip.paulis import int_to_limbs, symboltoint


def test_tfim_brick_32q_readme_value():
    # README.md:49-114 of the reference: 32 qubits, periodic brick TFIM, 32 layers, dt 0.1,
    # max_weight 6, min_abs_coeff 1e-4 -> overlapwithzero 0.154596728241...
    n = 32
    circ = pp.tfitrottercircuit(n, 32, topology=pp.bricklayertopology(n, periodic=True))
    th = np.full(pp.countparameters(circ), 0.1)
    obs = pp.PauliSum.from_string(pp.PauliString(n, "Z", 16))
    s, _ = oracle_propagate(circ, n, th, obs.terms, max_weight=6, min_abs_coeff=1e-4)
    assert abs(s.overlap_zero()[0] - 0.154596728241) < 1e-12


def test_tfim_6x6_notebook():
    # examples/advanced_performance.ipynb:92-118,182,224,438: 1,036,142 terms,
    # <Z21 Z22> = 0.6696851623009762 (VectorPauliSum), 200,640,660 gate applications
    nq = 36
    layer = []
    pp.rxlayer(layer, nq)
    pp.rzlayer(layer, nq)
    pp.rzzlayer(layer, pp.rectangletopology(6, 6))
    th = np.ones(len(layer)) * 0.05 * 2
    th[2 * nq:] *= 0.5
    obs = pp.PauliSum.from_string(pp.PauliString(nq, ["Z", "Z"], [21, 22]))
    s, counts = oracle_propagate(layer * 14, nq, np.tile(th, 14), obs.terms, min_abs_coeff=2.0 ** -20)
    assert len(s) == 1036142
    assert abs(s.overlap_zero()[0] - 0.6696851623009762) < 1e-14
    L = len(layer)
    assert sum(L * counts[i * L] for i in range(14)) == 200640660


def test_two_qubit_mcpropagate_example():
    # examples/mcpropagate.ipynb cell 3: [RY(1), RXX(1,2)], thetas [0.7, 0.5], Z1 -> 3 terms
    circ = [pp.PauliRotation("Y", 1), pp.PauliRotation(["X", "X"], [1, 2])]
    obs = pp.PauliSum.from_string(pp.PauliString(2, "Z", 1))
    s, _ = oracle_propagate(circ, 2, [0.7, 0.5], obs.terms, min_abs_coeff=0.0)
    keys, coeffs = s.get()
    got = {int(k[0]): float(c[0]) for k, c in zip(keys, coeffs)}
    YX = symboltoint(2, ["Y", "X"], [1, 2])
    ZI = symboltoint(2, "Z", 1)
    XI = symboltoint(2, "X", 1)
    assert set(got) == {YX, ZI, XI}
    assert abs(got[YX] - 0.47943) < 5e-6 and abs(got[ZI] - 0.67121) < 5e-6 and abs(got[XI] + 0.56535) < 5e-6
    assert s.overlap_zero()[0] == pytest.approx(0.6712121661589577, abs=1e-15)
    assert s.overlap_zero()[0] == pytest.approx(math.cos(0.5) * math.cos(0.7), abs=1e-15)


def test_tfim_3x3_layers():
    # examples/mcpropagate.ipynb cells 9-11: 105,012 terms after 5 layers and the rounded overlaps
    nq = 9
    layer = []
    pp.rxlayer(layer, nq)
    pp.rzlayer(layer, nq)
    pp.rzzlayer(layer, pp.rectangletopology(3, 3))
    th = np.ones(len(layer)) * 0.1 * 2
    th[2 * nq:] *= 0.5
    obs = pp.PauliSum.from_string(pp.PauliString(nq, ["Z", "Z"], [4, 5]))
    overlaps = [1.0]
    for nl in range(1, 6):
        s, _ = oracle_propagate(layer * nl, nq, np.tile(th, nl), obs.terms, min_abs_coeff=1e-8)
        overlaps.append(float(s.overlap_zero()[0]))
    assert len(s) == 105012
    assert [round(o, 4) for o in overlaps] == [1.0, 0.9605, 0.8592, 0.7366, 0.6333, 0.5723]


def test_merge_truncate_known_answer():
    # test/test_mergeandtruncate.jl:8-20: terms [1,1,2,2,3], coeffs [.60,-.55,.06,.06,1.0],
    # threshold 0.1 -> {2: 0.12, 3: 1.0} (0.60 - 0.55 cancels below the threshold)
    s = OracleSum(1, 1).load_merge(np.array([[1], [1], [2], [2], [3]], dtype=np.uint64),
                                   np.array([0.60, -0.55, 0.06, 0.06, 1.0]), min_abs=0.1)
    k, c = s.get()
    assert [int(x) for x in k[:, 0]] == [2, 3]
    assert c[0, 0] == 0.06 + 0.06 and c[1, 0] == 1.0


def test_xor_tail_sort_equals_comparison_sort():
    # test/test_xortailmerge.jl:30-91: the XOR passes give the comparison sort exactly
    rng = np.random.default_rng(0)
    for W, nq, masks in [(1, 32, [{1: 1}, {16: 3}, {3: 1, 9: 2}, {32: 3, 1: 3}]),
                         (2, 64, [{32: 3, 33: 3}, {5: 2, 60: 1}]),
                         (4, 127, [{63: 3, 65: 3}, {2: 1, 40: 2, 90: 3, 127: 1}])]:
        from helpers import random_terms
        terms = sorted(random_terms(nq, 5000, rng, 0.2))
        for codes in masks:
            m = sum(c << (2 * (q - 1)) for q, c in codes.items())
            tail = [t ^ m for t in terms]
            keys = np.array([int_to_limbs(t, W) for t in tail], dtype=np.uint64)
            co = np.arange(len(tail), dtype=np.float64).reshape(-1, 1)
            ml = np.array(int_to_limbs(m, W), dtype=np.uint64)
            orc.lib().ppo_xor_sort_tail(orc.u64p(keys), orc.dp(co), len(tail), W, 1, orc.u64p(ml))
            got = [sum(int(w) << (64 * l) for l, w in enumerate(row)) for row in keys]
            assert got == sorted(tail)
            # coefficients travelled with their keys
            assert [tail[int(i)] for i in co[:, 0]] == got


def test_overlap_is_sum_over_IZ_strings():
    rng = np.random.default_rng(3)
    from helpers import random_terms, terms_to_arrays
    terms = random_terms(40, 2000, rng, 0.1)
    keys, c = terms_to_arrays(terms, 40, rng.standard_normal(len(terms)))
    s = OracleSum(2, 1).set(keys, c)
    k, cc = s.get()
    assert s.overlap_zero()[0] == pytest.approx(fsum_overlap(k, cc)[0], abs=1e-12)


@pytest.mark.parametrize("zbits,threads", [(0, 1), (2, 4), (3, 3)])
def test_zone_parallel_baseline_equals_single_sum(zbits, threads):
    # the MultiPauliSum restatement (zones by a GF(2)-linear hash, tails sent to
    # zone ^ zone(mask)) gives the single sum's result bit for bit
    import ctypes
    from oracle import MultiOracleSum, OracleSum, lib
    from pauliprop_hip.gates import pack_circuit, toheisenberg
    nq = 24
    circ = []
    for _ in range(3):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, pp.bricklayertopology(nq, periodic=True))
        pp.depolarizinglayer(circ, nq, 0.02)
    P = pp.countparameters(circ)
    th = np.random.default_rng(zbits).uniform(0, 1.2, size=P)
    hc, hth = toheisenberg(circ, list(th))
    gates, _, _, _ = pack_circuit(hc, nq, 1)
    obs = {pp.symboltoint(nq, "Z", q): 1.0 for q in (5, 12, 20)}
    keys = np.array([[k] for k in obs], dtype=np.uint64)
    coeffs = np.ones((len(obs), 1))
    for mw, mac in ((6, 0.0), (8, 1e-4)):
        single = OracleSum(1, 1).set(keys, coeffs)
        c1 = np.zeros(len(hc), dtype=np.int64)
        lib().ppo_run(single._s, ctypes.cast(gates, ctypes.c_void_p), len(hc), np.ascontiguousarray(hth).ctypes.data_as(
            ctypes.POINTER(ctypes.c_double)), P, None, None, -1 if mw == math.inf else mw, mac,
            c1.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        multi = MultiOracleSum(1, 1, zbits).set(keys, coeffs)
        c2 = np.zeros(len(hc), dtype=np.int64)
        multi.run(ctypes.cast(gates, ctypes.c_void_p), len(hc), np.array(hth), P, mw, mac, threads, c2)
        assert np.array_equal(c1, c2)
        k1, v1 = single.get_sorted()
        k2, v2 = multi.get_sorted()
        assert np.array_equal(k1, k2) and np.array_equal(v1.view(np.uint64), v2.view(np.uint64))
