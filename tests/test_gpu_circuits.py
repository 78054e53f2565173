"""GPU parity, whole circuits: propagate() on the GPU against the oracle
(bit for bit) and against the reference's published values."""
import math

import numpy as np
import pytest

from helpers import assert_bitwise, fsum_overlap, rel_close
from oracle import oracle_propagate
import pauliprop_hip as pp

pytestmark = pytest.mark.gpu


def _tfim_brick_32q(nlayers=32):
    n = 32
    circ = pp.tfitrottercircuit(n, nlayers, topology=pp.bricklayertopology(n, periodic=True))
    th = np.full(pp.countparameters(circ), 0.1)
    obs = pp.PauliSum.from_string(pp.PauliString(n, "Z", 16))
    return n, circ, th, obs


def test_tfim_brick_32q_readme_value():
    # README.md:114 of the reference: overlapwithzero -> 0.154596728241...
    n, circ, th, obs = _tfim_brick_32q()
    dev = pp.propagate(circ, obs, th, max_weight=6, min_abs_coeff=1e-4)
    ref, _ = oracle_propagate(circ, n, th, obs.terms, max_weight=6, min_abs_coeff=1e-4)
    assert_bitwise(*dev.download(), *ref.get(), "32-qubit TFIM")
    ov = dev.overlapwithzero()
    assert abs(ov - 0.154596728241) < 1e-12
    assert rel_close(ov, float(ref.overlap_zero()[0]))


def test_tfim_brick_32q_weight_only_counts():
    n, circ, th, obs = _tfim_brick_32q()
    dev, counts = pp.propagate(circ, obs, th, max_weight=6, min_abs_coeff=0.0, counts=True)
    ref, rcounts = oracle_propagate(circ, n, th, obs.terms, max_weight=6, min_abs_coeff=0.0)
    assert np.array_equal(counts, rcounts)
    assert_bitwise(*dev.download(), *ref.get(), "32-qubit TFIM weight-only")
    assert rel_close(dev.overlapwithzero(), fsum_overlap(*ref.get())[0])


def test_tfim_6x6_notebook_anchor():
    # examples/advanced_performance.ipynb:92-118,182,224 of the reference:
    # 1,036,142 terms and <Z21 Z22> = 0.6696851623009762
    nq = 36
    layer = []
    pp.rxlayer(layer, nq)
    pp.rzlayer(layer, nq)
    pp.rzzlayer(layer, pp.rectangletopology(6, 6))
    th = np.ones(len(layer)) * 0.05 * 2
    th[2 * nq:] *= 0.5
    circ, cth = layer * 14, np.tile(th, 14)
    obs = pp.PauliSum.from_string(pp.PauliString(nq, ["Z", "Z"], [21, 22]))
    dev = pp.propagate(circ, obs, cth, min_abs_coeff=2.0 ** -20)
    assert len(dev) == 1036142
    assert abs(dev.overlapwithzero() - 0.6696851623009762) < 1e-13


@pytest.mark.parametrize("B", [2, 5, 64])
def test_batched_heisenberg_matches_columns(B):
    # one shared key set, B parameter columns: weight-only truncation gives the
    # same keys for every column, and each column equals its own B=1 run
    nx = 3
    nq = nx * nx
    topo = pp.rectangletopology(nx, nx, periodic=True)
    circ = []
    for _ in range(2):
        pp.rxxlayer(circ, topo)
        pp.ryylayer(circ, topo)
        pp.rzzlayer(circ, topo)
        pp.rzlayer(circ, nq)
    P = pp.countparameters(circ)
    rng = np.random.default_rng(B)
    th = rng.uniform(0, 0.2, size=(P, B))
    obs = pp.PauliSum(nq)
    for a, b in topo:
        obs.add(["Z", "Z"], [a, b], 1.0)
    dev = pp.propagate(circ, obs, th, max_weight=5, min_abs_coeff=0.0)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=5, min_abs_coeff=0.0, batch=B)
    assert_bitwise(*dev.download(), *ref.get(), f"batched B={B}")
    ov = dev.overlapwithzero()
    exact = fsum_overlap(*ref.get())
    for b in range(B):
        assert rel_close(ov[b], exact[b])


def test_batched_coefficient_truncation_uses_batch_max():
    nq = 16
    circ = pp.tfitrottercircuit(nq, 4, topology=pp.bricklayertopology(nq, periodic=True))
    P = pp.countparameters(circ)
    B = 8
    th = np.random.default_rng(1).uniform(0, 0.6, size=(P, B))
    obs = pp.PauliSum.from_string(pp.PauliString(nq, "Z", 8))
    dev = pp.propagate(circ, obs, th, min_abs_coeff=1e-3)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, min_abs_coeff=1e-3, batch=B)
    assert_bitwise(*dev.download(), *ref.get(), "batched truncation")


def test_noisy_circuit():
    nq = 12
    circ = []
    for _ in range(3):
        pp.rylayer(circ, nq)
        pp.rzlayer(circ, nq)
        pp.rzzlayer(circ, pp.staircasetopology(nq))
        pp.depolarizinglayer(circ, nq, 0.01)
    P = pp.countparameters(circ)
    th = np.random.default_rng(5).normal(0, 0.3, size=P)
    obs = pp.PauliSum(nq)
    for i in range(1, nq):
        obs.add(["Z", "Z"], [i, i + 1])
    dev = pp.propagate(circ, obs, th, max_weight=7, min_abs_coeff=1e-5)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=7, min_abs_coeff=1e-5)
    assert_bitwise(*dev.download(), *ref.get(), "noisy")


@pytest.mark.parametrize("B", [1, 3, 64])
@pytest.mark.parametrize("mode", ["default", "tiny_spin_limit"])
def test_listed_touched_terms_match_own_pass(mode, B):
    # a merge lists the next rotation's touched terms; the result must equal the
    # run where every rotation lists them itself, and the oracle, bit for bit,
    # also when gates are stopped and run again (one-poll look-back waits) and
    # when noise gates break the chain of rotations
    print(f"listed touched terms, {mode}, B={B}", flush=True)
    nq = 40
    circ = []
    for layer in range(4):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, pp.bricklayertopology(nq, periodic=True))
        if layer == 1:
            pp.depolarizinglayer(circ, nq, 0.02)
    P = pp.countparameters(circ)
    th = np.random.default_rng(9).uniform(0.1, 1.2, size=(P, B))
    if B == 1:
        th = th[:, 0]
    obs = pp.PauliSum(nq)
    for q in (7, 20, 33):
        obs.add(["Z"], [q])
    results = []
    for chain in (True, False):
        dev = pp.HIPPauliSum.from_paulisum(obs, batch=B)
        dev.tune(chain_tails=chain, lds_runs=False)  # the per-gate pipeline's chaining
        if mode == "tiny_spin_limit":
            dev.tune(spin_limit=1)
        pp.propagate(circ, dev, th, max_weight=8, min_abs_coeff=1e-6)
        results.append(dev.download())
    assert_bitwise(*results[0], *results[1], f"chained vs own pass ({mode})")
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=8, min_abs_coeff=1e-6, batch=B)
    assert_bitwise(*results[0], *ref.get(), f"chained vs oracle ({mode})")


@pytest.mark.parametrize("B", [1, 2])
def test_identity_rotations_left_out(B):
    # a light cone: the observable sits on two qubits, so most early rotations
    # touch nothing and are left out (per-qubit Pauli presence read every 64
    # gates, widened over rotations, noise and Cliffords in between); the sum,
    # every gate's input size and the oracle's must agree bit for bit
    print(f"identity rotations left out, B={B}", flush=True)
    nq = 60
    circ = []
    topo = pp.bricklayertopology(nq)
    for layer in range(6):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, topo)
        if layer == 2:
            pp.depolarizinglayer(circ, nq, 0.01)
        if layer == 3:
            for q in range(20, 40, 2):
                circ.append(pp.CliffordGate("CNOT", [q, q + 1]))
            circ.append(pp.CliffordGate("H", [31]))
    P = pp.countparameters(circ)
    th = np.random.default_rng(17).uniform(0.2, 1.3, size=(P, B))
    obs = pp.PauliSum(nq)
    obs.add(["Z"], [30])
    obs.add(["Z", "Z"], [29, 31], 0.5)
    thb = th if B > 1 else th[:, 0]
    runs = []
    for skip in (True, False):
        dev = pp.HIPPauliSum.from_paulisum(obs, batch=B)
        dev.tune(skip_identity=skip, lds_runs=False)  # the per-gate pipeline's presence test
        dev.profile_control(reset=True)
        _, cnt = pp.propagate(circ, dev, thb, max_weight=7, min_abs_coeff=1e-6, counts=True)
        runs.append((dev.download(), np.array(cnt), dev.profile()["identity_gates"]))
    (k1, c1), n1, skipped = runs[0]
    (k0, c0), n0, none = runs[1]
    print(f"left out {skipped} of {len(circ)} gates", flush=True)
    assert skipped > len(circ) // 4 and none == 0
    assert_bitwise(k1, c1, k0, c0, f"left out vs all gates, B={B}")
    assert np.array_equal(n1, n0)
    ref, rcnt = oracle_propagate(circ, nq, thb, obs.terms, max_weight=7, min_abs_coeff=1e-6, batch=B)
    assert_bitwise(k1, c1, *ref.get_sorted(), f"left out vs oracle, B={B}")
    assert np.array_equal(n1, np.asarray(rcnt))


@pytest.mark.parametrize("B", [1, 3])
def test_amplitude_damping_circuit(B):
    print(f"amplitude damping circuit B={B}", flush=True)
    nq = 14
    circ = []
    for _ in range(3):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, pp.bricklayertopology(nq))
        for q in range(1, nq + 1, 3):
            circ.append(pp.FrozenGate(pp.AmplitudeDampingNoise(q), 0.05))
    P = pp.countparameters(circ)
    th = np.random.default_rng(12).uniform(-1, 1, size=(P, B))
    obs = pp.PauliSum(nq)
    for q in (3, 8):
        obs.add(["Z"], [q])
    dev = pp.propagate(circ, obs, th if B > 1 else th[:, 0], max_weight=6, min_abs_coeff=1e-7)
    ref, _ = oracle_propagate(circ, nq, th if B > 1 else th[:, 0], obs.terms, max_weight=6, min_abs_coeff=1e-7,
                              batch=B)
    assert_bitwise(*dev.download(), *ref.get(), f"amplitude damping circuit B={B}")


@pytest.mark.parametrize("B", [1, 2])
def test_clifford_circuit(B):
    # rotations, CNOT/H/S layers and noise in one gate list (pp_run_circuit with tables)
    print(f"Clifford circuit B={B}", flush=True)
    nq = 20
    circ = []
    for layer in range(3):
        pp.rylayer(circ, nq)
        for q in range(1, nq, 2):
            circ.append(pp.CliffordGate("CNOT", [q, q + 1]))
        circ.append(pp.CliffordGate("H", [layer + 1]))
        circ.append(pp.CliffordGate("S", [nq - layer]))
        pp.rzlayer(circ, nq)
        pp.depolarizinglayer(circ, nq, 0.01)
    circ.append(pp.CliffordGate("CZ", [2, 19]))  # the last gate leaves the oracle unsorted
    P = pp.countparameters(circ)
    th = np.random.default_rng(31).uniform(-1, 1, size=(P, B))
    obs = pp.PauliSum(nq)
    for q in (4, 11, 17):
        obs.add(["Z"], [q])
    for mw in (math.inf, 6):
        dev = pp.propagate(circ, obs, th if B > 1 else th[:, 0], max_weight=mw, min_abs_coeff=1e-6)
        ref, _ = oracle_propagate(circ, nq, th if B > 1 else th[:, 0], obs.terms, max_weight=mw,
                                  min_abs_coeff=1e-6, batch=B)
        assert_bitwise(*dev.download(), *ref.get_sorted(), f"Clifford circuit B={B} max_weight={mw}")


@pytest.mark.parametrize("B", [65, 128, 300])
def test_wide_batches_rotation_and_noise_bitwise(B):
    # the batched merge's and noise pass's loops over columns past 64 and the
    # 32- and 16-row merge tiles that batches above 64 take
    print(f"wide batch B={B}", flush=True)
    nq = 12
    circ = []
    for _ in range(2):
        pp.rylayer(circ, nq)
        pp.rzzlayer(circ, pp.bricklayertopology(nq))
        pp.depolarizinglayer(circ, nq, 0.05)
        pp.rxlayer(circ, nq)
    P = pp.countparameters(circ)
    th = np.random.default_rng(B).uniform(-0.8, 0.8, size=(P, B))
    obs = pp.PauliSum(nq)
    for i in range(1, nq):
        obs.add(["Z", "Z"], [i, i + 1], 1.0)
    dev = pp.propagate(circ, obs, th, max_weight=6, min_abs_coeff=1e-4)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=6, min_abs_coeff=1e-4, batch=B)
    assert_bitwise(*dev.download(), *ref.get(), f"wide batch B={B}")


def test_wide_batch_amplitude_damping_bitwise():
    print("wide batch amplitude damping B=65", flush=True)
    nq, B = 10, 65
    circ = []
    for q in range(1, nq + 1):
        circ.append(pp.PauliRotation("X", q))
        circ.append(pp.freeze(pp.AmplitudeDampingNoise(q), 0.1 + 0.02 * q))
    pp.rzzlayer(circ, pp.bricklayertopology(nq))
    for q in range(1, nq + 1, 2):
        circ.append(pp.freeze(pp.AmplitudeDampingNoise(q), 0.3))
    P = pp.countparameters(circ)
    th = np.random.default_rng(7).uniform(-1, 1, size=(P, B))
    obs = pp.PauliSum(nq)
    for i in range(1, nq):
        obs.add(["Z", "Z"], [i, i + 1], 1.0)
    dev = pp.propagate(circ, obs, th, max_weight=5, min_abs_coeff=1e-5)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=5, min_abs_coeff=1e-5, batch=B)
    assert_bitwise(*dev.download(), *ref.get(), "wide batch amplitude damping")


@pytest.mark.timeout(600)
def test_heisenberg_64q_b64_first_step_matches_zone_oracle():
    # the batched workload (8x8 periodic Heisenberg, 64 qubits, UInt128 keys,
    # B = 64 columns, max_weight 8, min_abs_coeff 1e-6 on the batch maximum) at
    # its own size for its first Trotter step (448 gates, 1,455,699 terms):
    # bit for bit against the zone-parallel oracle on the host's cores
    import ctypes
    import os
    from oracle import MultiOracleSum
    from pauliprop_hip import workloads
    from pauliprop_hip.hipsum import PackedCircuit
    print("64-qubit B=64 Heisenberg, first step", flush=True)
    wl = workloads.heisenberg_square_64q_b64(nsteps=1)
    B = wl["batch"]
    pc = PackedCircuit(wl["circuit"], wl["thetas"], 64, batch=B)
    keys0, coeffs0 = wl["obs"].to_arrays(pc.W, B)
    dev = pp.HIPPauliSum(64, batch=B, capacity=1 << 20)
    dev.upload(keys0, coeffs0)
    counts = pc.run(dev, max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"]).copy()
    print(f"device: {len(dev)} terms", flush=True)
    m = MultiOracleSum(pc.W, B, zbits=4)
    m.set(keys0, coeffs0)
    rcounts = np.zeros(pc.ngates, dtype=np.int64)
    m.run(ctypes.cast(pc.gates, ctypes.c_void_p), pc.ngates, pc.theta, pc.nparams, wl["max_weight"],
          wl["min_abs_coeff"], min(16, os.cpu_count() or 1), rcounts)
    assert np.array_equal(counts, rcounts)
    assert len(dev) == 1455699
    assert_bitwise(*dev.download(), *m.get_sorted(), "64-qubit B=64 first step")
