"""GPU parity of the two gradients.

pp_adjoint (pp.adjointgradient): the fixed-truncation adjoint sweep, now with
Clifford gates, the other linear overlaps (overlapwithplus, computational,
maxmixed), batches wider than a workgroup, and the 48-qubit noisy B = 32
ansatz workload, each against the oracle's reverse sweep.

pp_rewind (pp.rewindgradient): the reference's own gradient algorithm
(src/gradient.jl:40-108) against the oracle's restatement of it, under
truncation, on the reference's test circuits (test/test_gradient.jl) and on the
48-qubit ansatz workload; and equal to pp_adjoint when nothing is truncated."""
import math

import numpy as np
import pytest

from oracle import oracle_adjoint, oracle_propagate, oracle_rewind
import pauliprop_hip as pp
from pauliprop_hip import workloads

pytestmark = pytest.mark.gpu

R, C, F = pp.PauliRotation, pp.CliffordGate, pp.freeze


def _close(got, want, tol=1e-12):
    scale = max(1.0, float(np.abs(want).max()) if np.size(want) else 1.0)
    np.testing.assert_allclose(got, want, rtol=0, atol=tol * scale)


def _clifford_ansatz(n, layers, noise):
    circ = []
    for k in range(layers):
        pp.rylayer(circ, n)
        circ.append(C(["H", "SX", "S", "SY"][k % 4], 1 + k % n))
        pp.rzlayer(circ, n)
        circ.append(C(["CNOT", "CZ", "ZZPIHALF", "SWAP"][k % 4], [1 + k % n, 1 + (k + 3) % n]))
        pp.rzzlayer(circ, pp.staircasetopology(n))
        if noise:
            pp.depolarizinglayer(circ, n, 0.02)
    obs = pp.PauliSum(n)
    for i in range(1, n):
        obs.add(["Z", "Z"], [i, i + 1])
    obs.add(["X"], [2], 0.5)
    return circ, obs


@pytest.mark.parametrize("overlap,ones", [("zero", ()), ("plus", ()), ("computational", (1, 4, 5)),
                                          ("maxmixed", ())])
@pytest.mark.parametrize("noise", [False, True])
def test_adjoint_with_cliffords_and_overlaps_matches_oracle(overlap, ones, noise):
    print(f"adjoint cliffords overlap={overlap} noise={noise}", flush=True)
    n = 8
    circ, obs = _clifford_ansatz(n, 3, noise)
    P = pp.countparameters(circ)
    th = np.random.default_rng(40 + noise).normal(0, 0.4, size=P)
    loss, grad = pp.adjointgradient(circ, obs, th, max_weight=5, min_abs_coeff=1e-5, overlapfunc=overlap,
                                    onebitinds=ones)
    rloss, rgrad = oracle_adjoint(circ, n, th, obs.terms, max_weight=5, min_abs_coeff=1e-5, overlap=overlap,
                                  onebitinds=ones)
    assert loss == pytest.approx(rloss[0], rel=1e-12, abs=1e-14)
    _close(grad, rgrad[:, 0])
    if overlap == "maxmixed":
        assert np.all(np.abs(grad) < 1e-14)  # the trace is conserved by unitaries and Pauli noise


@pytest.mark.parametrize("B", [257, 300])
def test_adjoint_batches_wider_than_a_workgroup(B):
    # columns beyond 256 are covered by column groups of the sweep's rows
    print(f"adjoint B={B}", flush=True)
    n = 6
    circ, obs = _clifford_ansatz(n, 2, True)
    P = pp.countparameters(circ)
    th = np.random.default_rng(B).normal(0, 0.3, size=(P, B))
    loss, grad = pp.adjointgradient(circ, obs, th, max_weight=4, min_abs_coeff=1e-6)
    rloss, rgrad = oracle_adjoint(circ, n, th, obs.terms, max_weight=4, min_abs_coeff=1e-6, batch=B)
    np.testing.assert_allclose(loss, rloss, rtol=1e-12, atol=1e-14)
    _close(grad, rgrad)


def test_adjoint_deferred_and_per_gate_sums_are_identical():
    print("adjoint partial-gradient store on and off", flush=True)
    n = 9
    circ, obs = _clifford_ansatz(n, 3, True)
    P = pp.countparameters(circ)
    th = np.random.default_rng(9).normal(0, 0.3, size=(P, 3))
    pc = pp.PackedCircuit(circ, th, n, batch=3)
    a = pp.HIPPauliSum.from_paulisum(obs, batch=3)
    l0, g0 = pc.adjoint(a, max_weight=6, min_abs_coeff=1e-6)
    b = pp.HIPPauliSum.from_paulisum(obs, batch=3)
    b.tune(adj_defer_bytes=0)
    l1, g1 = pc.adjoint(b, max_weight=6, min_abs_coeff=1e-6)
    assert np.array_equal(l0, l1) and np.array_equal(g0.view(np.uint64), g1.view(np.uint64))


@pytest.mark.timeout(900)
def test_noisy_48q_b32_ansatz_matches_oracle_and_differences():
    # the 48-qubit ansatz with depolarizing noise after every layer, B = 32
    # columns, 1716 parameters: the whole batched adjoint against the oracle,
    # and central differences (h = 1e-6) of device forward runs on 8
    # parameters across two columns
    print("48-qubit noisy B=32 adjoint", flush=True)
    wl = workloads.ansatz_line_noisy_48q_b32()
    n, mw, mac = wl["nqubits"], wl["max_weight"], wl["min_abs_coeff"]
    th = wl["thetas"]
    loss, grad = pp.adjointgradient(wl["circuit"], wl["obs"], th, max_weight=mw, min_abs_coeff=mac)
    print("device done; oracle running", flush=True)
    rloss, rgrad = oracle_adjoint(wl["circuit"], n, th, wl["obs"].terms, max_weight=mw, min_abs_coeff=mac,
                                  batch=32)
    np.testing.assert_allclose(loss, rloss, rtol=1e-12)
    _close(grad, rgrad, tol=1e-11)
    pc = pp.PackedCircuit(wl["circuit"], th, n, batch=32)
    keys0, coeffs0 = wl["obs"].to_arrays(pc.W, 32)
    psum = pp.HIPPauliSum(n, batch=32, capacity=1 << 20)

    def loss_at(t):
        pc.theta[:pc.nparams] = t[::-1]  # Heisenberg rows
        psum.upload(keys0, coeffs0)
        pc.run(psum, max_weight=mw, min_abs_coeff=mac)
        return psum.overlapwithzero()

    eps = 1e-6
    worst = 0.0
    for b in (0, 17):
        for k in (0, 211, 700, 1715):
            tp, tm = th.copy(), th.copy()
            tp[k, b] += eps
            tm[k, b] -= eps
            fd = (loss_at(tp)[b] - loss_at(tm)[b]) / (2 * eps)
            worst = max(worst, abs(fd - grad[k, b]))
            assert abs(fd - grad[k, b]) < 1e-7, (b, k, fd, grad[k, b])
    print(f"max |adjoint - central difference| = {worst:.3e}", flush=True)


# ---------------------------------------------------------------- rewindgradient on the device

def _zz(nq, a, b, c=1.0):
    return {pp.symboltoint(nq, ["Z", "Z"], [a, b]): c}


REF_CASES = {
    # the circuits of the reference's test/test_gradient.jl
    "rotations": ([R("X", 1), R("Y", 2), R(["X", "Z"], [1, 3]), R("Z", 4), R(["Y", "Y"], [2, 4])],
                  4, [0.3, -0.7, 1.1, 0.42, -0.15], _zz(4, 1, 3)),
    "cliffords": ([R("X", 1), C("H", 2), R(["Y", "Z"], [1, 3]), C("CNOT", [2, 3]), R("Z", 4), C("S", 4),
                   R(["X", "X"], [1, 4])], 4, [0.2, -0.5, 0.83, -1.0], _zz(4, 1, 4)),
    "frozen": ([R("X", 1), F(R("Y", 2), 0.66), R(["X", "Z"], [1, 3]), F(R("Z", 4), -0.31), R(["Y", "Y"], [2, 4])],
               4, [0.44, -0.9, 0.12], _zz(4, 1, 3)),
    "mixed": ([R("X", 1), C("H", 2), F(R("Y", 2), 0.5), C("CZ", [2, 3]), R(["X", "Z"], [1, 3]), C("SX", 4),
               F(R("Z", 4), -0.2), R(["Y", "X"], [3, 5]), C("CNOT", [4, 5]), R("X", 5)],
              5, [0.31, -0.62, 0.9, -0.44], {pp.symboltoint(5, "Z", 3): 1.0}),
}


@pytest.mark.parametrize("case", sorted(REF_CASES))
@pytest.mark.parametrize("overlap,ones", [("zero", ()), ("plus", ()), ("computational", (1, 3)), ("maxmixed", ())])
@pytest.mark.parametrize("trunc", [(math.inf, 0.0), (2, 0.05)])
def test_rewind_matches_oracle_rewind(case, overlap, ones, trunc):
    circ, nq, params, obs = REF_CASES[case]
    mw, mac = trunc
    loss, grad = pp.rewindgradient(circ, pp.PauliSum(nq, obs), params, overlapfunc=overlap,
                                   onebitinds=ones, max_weight=mw, min_abs_coeff=mac)
    rloss, rgrad = oracle_rewind(circ, nq, params, obs, max_weight=mw, min_abs_coeff=mac, overlap=overlap,
                                 onebitinds=ones)
    assert loss == pytest.approx(rloss, rel=1e-12, abs=1e-15)
    _close(grad, rgrad)


def test_rewind_rejects_noise_and_batches():
    circ = [R("X", 1), pp.DepolarizingNoise(2, 0.1), R("Y", 2)]
    with pytest.raises(pp._lib.PPError):
        pp.rewindgradient(circ, pp.PauliSum(3, _zz(3, 1, 2)), [0.3, 0.4])


def test_rewind_truncation_to_empty():
    circ = [R("X", 1), R("Z", 2), R("Y", 1)]
    loss, g = pp.rewindgradient(circ, pp.PauliSum(2, {pp.symboltoint(2, "Z", 1): 1.0}), [0.3, 0.7, 0.2],
                                min_abs_coeff=10.0)
    assert loss == 0.0 and np.all(g == 0)


def test_adjoint_equals_rewind_without_truncation():
    # the 48-qubit ansatz's gates on 12 qubits, nothing truncated: both
    # gradients are the exact derivative
    print("adjoint == rewind without truncation", flush=True)
    n = 12
    circ, obs = [], pp.PauliSum(n)
    for _ in range(2):
        pp.rylayer(circ, n)
        pp.rzlayer(circ, n)
        pp.rzzlayer(circ, pp.staircasetopology(n))
    for i in range(1, n):
        obs.add(["Z", "Z"], [i, i + 1])
    P = pp.countparameters(circ)
    th = workloads.SplitMix64(4801).normals(P, 0.3)
    l1, g1 = pp.adjointgradient(circ, obs, th, min_abs_coeff=0.0)
    l2, g2 = pp.rewindgradient(circ, obs, th, min_abs_coeff=0.0)
    assert l1 == pytest.approx(l2, rel=1e-12)
    _close(g1, g2)


def test_rewind_48q_ansatz_matches_oracle_rewind():
    # the 48-qubit ansatz workload at its own truncation (max_weight 7,
    # min_abs_coeff 1e-5): the device runs the reference's algorithm step for
    # step, so its gradient equals the restatement's; the fixed-truncation
    # adjoint differs from it by the truncation's effect, reported here
    print("48-qubit ansatz rewindgradient", flush=True)
    wl = workloads.ansatz_line_48q()
    n, mw, mac = wl["nqubits"], wl["max_weight"], wl["min_abs_coeff"]
    loss, grad = pp.rewindgradient(wl["circuit"], wl["obs"], wl["thetas"], max_weight=mw, min_abs_coeff=mac)
    rloss, rgrad = oracle_rewind(wl["circuit"], n, wl["thetas"], wl["obs"].terms, max_weight=mw, min_abs_coeff=mac)
    assert loss == pytest.approx(rloss, rel=1e-12)
    _close(grad, rgrad, tol=1e-11)
    aloss, agrad = pp.adjointgradient(wl["circuit"], wl["obs"], wl["thetas"], max_weight=mw, min_abs_coeff=mac)
    assert aloss == pytest.approx(loss, rel=1e-12)
    dev = np.abs(agrad - grad)
    print(f"|adjoint - rewindgradient|: max {dev.max():.3e}, rms ratio "
          f"{math.sqrt((dev ** 2).mean() / (grad ** 2).mean()):.3e}, max|grad| {np.abs(grad).max():.3e}", flush=True)
    assert dev.max() < 1e-2 * np.abs(grad).max()
