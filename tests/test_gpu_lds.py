"""GPU parity of the shared-memory runs (k_lds_run): circuits whose sums are
small enough for one workgroup to hold run many gates per launch. Every run is
compared bit for bit with the same circuit on the per-gate pipeline
(PP_TUNE_LDS_RUNS off) and with the oracle, and the profile shows which path
applied how many gates."""
import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.hipsum import HIPPauliSum

pytestmark = pytest.mark.gpu


def _run(circ, obs, th, lds, capacity=1 << 14, **kw):
    s = HIPPauliSum.from_paulisum(obs, capacity=capacity)
    s.tune(lds_runs=lds)
    s.profile_control(reset=True)
    s, counts = pp.propagate(circ, s, th, counts=True, **kw)
    return s, counts, s.profile()


def _check(circ, nq, obs, th, what, expect_lds=True, expect_pipeline=None, capacity=1 << 14, **kw):
    on, c_on, p_on = _run(circ, obs, th, True, capacity=capacity, **kw)
    off, c_off, p_off = _run(circ, obs, th, False, capacity=capacity, **kw)
    ref, rcounts = oracle_propagate(circ, nq, th, obs.terms, **kw)
    print(what, "terms", len(on), "persistent", p_on["persistent_gates"], "merges",
          p_on["kernels"]["merge"]["launches"], flush=True)
    assert np.array_equal(c_on, c_off) and np.array_equal(c_on, rcounts), f"{what}: gate sizes differ"
    assert_bitwise(*on.download(), *off.download(), f"{what}: shared-memory run vs per-gate pipeline")
    assert_bitwise(*on.download(), *ref.get(), f"{what}: vs oracle")
    assert p_off["persistent_gates"] == 0
    if expect_lds:
        assert p_on["persistent_gates"] > 0, f"{what}: no gate ran in shared memory"
    if expect_pipeline is not None:
        assert (p_on["kernels"]["merge"]["launches"] > 0) == expect_pipeline, f"{what}: pipeline use"
    return on


def test_tfim_brick_32q_in_shared_memory():
    # the README workload (at most 1671 terms): every gate in shared-memory runs
    n = 32
    circ = pp.tfitrottercircuit(n, 32, topology=pp.bricklayertopology(n, periodic=True))
    th = np.full(pp.countparameters(circ), 0.1)
    obs = pp.PauliSum.from_string(pp.PauliString(n, "Z", 16))
    on = _check(circ, n, obs, th, "tfim 32q", expect_pipeline=False, max_weight=6, min_abs_coeff=1e-4)
    assert abs(on.overlapwithzero() - 0.154596728241) < 1e-12


def _noisy_layers(nq, nlayers, rng, p_noise):
    """RX on every qubit, RZZ on a line, RY on every other qubit, then
    depolarizing noise on every qubit: the sum grows past shared memory and the
    noise and truncation bring it back, so runs stop and start again."""
    circ = []
    for layer in range(nlayers):
        pp.rxlayer(circ, nq)
        for q in range(1, nq):
            circ.append(pp.PauliRotation(["Z", "Z"], [q, q + 1]))
        for q in range(1, nq + 1, 2):
            circ.append(pp.PauliRotation(["Y"], [q]))
        if layer % 2 == 1:
            for q in range(1, nq + 1):
                circ.append(pp.DepolarizingNoise(q, p_noise))
    th = rng.uniform(0.1, 0.6, pp.countparameters(circ))
    return circ, th


@pytest.mark.parametrize("nq", [20, 40, 100])
def test_runs_stop_and_resume(nq):
    rng = np.random.default_rng(nq)
    circ, th = _noisy_layers(nq, 6, rng, 0.25)
    obs = pp.PauliSum(nq)
    obs.add(["Z"], [nq // 2], 1.0)
    obs.add(["X", "X"], [1, 2], 0.5)
    _check(circ, nq, obs, th, f"noisy {nq}q", max_weight=5, min_abs_coeff=2e-3)


@pytest.mark.parametrize("nq,mw", [(24, 4), (60, 5), (127, 4)])
def test_weight_only_truncation(nq, mw):
    # no coefficient threshold: sets and sizes fixed by the weight limit alone
    rng = np.random.default_rng(7 * nq)
    circ, th = _noisy_layers(nq, 3, rng, 0.1)
    obs = pp.PauliSum.from_string(pp.PauliString(nq, "Z", 1))
    _check(circ, nq, obs, th, f"weight-only {nq}q", max_weight=mw, min_abs_coeff=0.0)


def test_small_capacity_stops_runs():
    # a handle smaller than shared memory: a run stops where the sum outgrows
    # the handle, the pipeline grows the buffers and goes on
    nq = 16
    rng = np.random.default_rng(3)
    circ, th = _noisy_layers(nq, 4, rng, 0.05)
    obs = pp.PauliSum.from_string(pp.PauliString(nq, "Z", 8))
    _check(circ, nq, obs, th, "capacity 64", capacity=64, max_weight=6, min_abs_coeff=1e-4)


def test_kicked_ising_first_layers():
    # the 127-qubit kicked Ising workload's first two layers start in shared
    # memory and leave it as the sum grows
    from pauliprop_hip import workloads
    wl = workloads.kicked_ising_heavyhex_127q()
    circ = wl["circuit"][-2 * 271:]
    th = wl["thetas"][-2 * 271:]
    _check(circ, 127, wl["obs"], th, "kicked Ising 2 layers", max_weight=wl["max_weight"],
           min_abs_coeff=wl["min_abs_coeff"])
