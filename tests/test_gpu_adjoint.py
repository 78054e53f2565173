"""GPU parity of the adjoint sweep (pp_adjoint): against the oracle's reverse
sweep, against central finite differences of the GPU forward run
(test/test_gradient.jl:9-31 of the reference: eps 1e-6, atol 1e-7), and for
checkpoint segments, batches and the 48-qubit ansatz workload."""
import math

import numpy as np
import pytest

from oracle import oracle_adjoint
import pauliprop_hip as pp
from pauliprop_hip import workloads

pytestmark = pytest.mark.gpu


def _ansatz(n, layers, noise):
    circ = []
    for _ in range(layers):
        pp.rylayer(circ, n)
        pp.rzlayer(circ, n)
        pp.rzzlayer(circ, pp.staircasetopology(n))
        if noise:
            pp.depolarizinglayer(circ, n, 0.02)
    obs = pp.PauliSum(n)
    for i in range(1, n):
        obs.add(["Z", "Z"], [i, i + 1])
    return circ, obs


def _gpu_loss(circ, obs, th, **kw):
    return pp.propagate(circ, obs, th, **kw).overlapwithzero()


@pytest.mark.parametrize("noise", [False, True])
def test_adjoint_matches_oracle_and_central_differences(noise):
    print(f"adjoint vs oracle and FD, noise={noise}", flush=True)
    n = 8
    circ, obs = _ansatz(n, 3, noise)
    P = pp.countparameters(circ)
    th = np.random.default_rng(3 + noise).normal(0, 0.3, size=P)
    loss, grad = pp.adjointgradient(circ, obs, th, max_weight=5, min_abs_coeff=0.0)
    rloss, rgrad = oracle_adjoint(circ, n, th, obs.terms, max_weight=5, min_abs_coeff=0.0)
    assert loss == pytest.approx(rloss[0], rel=1e-12, abs=1e-14)
    scale = max(1.0, float(np.abs(rgrad).max()))
    np.testing.assert_allclose(grad, rgrad[:, 0], rtol=0, atol=1e-12 * scale)
    # weight truncation does not depend on theta, so central differences see the same masks
    eps = 1e-6
    for k in range(0, P, 5):
        tp, tm = th.copy(), th.copy()
        tp[k] += eps
        tm[k] -= eps
        fd = (_gpu_loss(circ, obs, tp, max_weight=5, min_abs_coeff=0.0) -
              _gpu_loss(circ, obs, tm, max_weight=5, min_abs_coeff=0.0)) / (2 * eps)
        assert abs(fd - grad[k]) < 1e-7, (k, fd, grad[k])


def test_adjoint_with_coefficient_truncation_matches_oracle():
    # truncation by |c| is held as the forward run made it, in both implementations
    print("adjoint with min_abs_coeff", flush=True)
    n = 10
    circ, obs = _ansatz(n, 4, True)
    P = pp.countparameters(circ)
    th = np.random.default_rng(8).normal(0, 0.4, size=P)
    loss, grad = pp.adjointgradient(circ, obs, th, max_weight=6, min_abs_coeff=1e-4)
    rloss, rgrad = oracle_adjoint(circ, n, th, obs.terms, max_weight=6, min_abs_coeff=1e-4)
    assert loss == pytest.approx(rloss[0], rel=1e-12, abs=1e-14)
    scale = max(1.0, float(np.abs(rgrad).max()))
    np.testing.assert_allclose(grad, rgrad[:, 0], rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("segment", [1, 7, 40])
def test_checkpoint_segments_give_identical_gradients(segment):
    # each segment is run again from a snapshot of its input sum; the rerun is
    # bit-identical, so the gradient is too, and the handle ends on the final sum
    print(f"checkpoint segment {segment}", flush=True)
    n = 9
    circ, obs = _ansatz(n, 3, True)
    P = pp.countparameters(circ)
    th = np.random.default_rng(21).normal(0, 0.3, size=P)
    pc = pp.PackedCircuit(circ, th, n)
    whole = pp.HIPPauliSum.from_paulisum(obs)
    l0, g0 = pc.adjoint(whole, max_weight=6, min_abs_coeff=1e-6)
    seg = pp.HIPPauliSum.from_paulisum(obs)
    l1, g1 = pc.adjoint(seg, max_weight=6, min_abs_coeff=1e-6, checkpoint_gates=segment)
    assert np.array_equal(l0, l1)
    assert np.array_equal(g0, g1)
    k0, c0 = whole.download()
    k1, c1 = seg.download()
    assert np.array_equal(k0, k1) and np.array_equal(c0.view(np.uint64), c1.view(np.uint64))


def test_batched_adjoint_matches_oracle():
    print("batched adjoint", flush=True)
    n = 8
    circ, obs = _ansatz(n, 3, True)
    P = pp.countparameters(circ)
    B = 5
    th = np.random.default_rng(4).normal(0, 0.3, size=(P, B))
    loss, grad = pp.adjointgradient(circ, obs, th, max_weight=5, min_abs_coeff=1e-5)
    rloss, rgrad = oracle_adjoint(circ, n, th, obs.terms, max_weight=5, min_abs_coeff=1e-5, batch=B)
    np.testing.assert_allclose(loss, rloss, rtol=1e-12, atol=1e-14)
    scale = max(1.0, float(np.abs(rgrad).max()))
    np.testing.assert_allclose(grad, rgrad, rtol=0, atol=1e-12 * scale)


def test_ansatz_48_qubits_matches_oracle():
    # the 48-qubit, 12-layer ansatz workload: about 1700 parameters, 67k peak terms;
    # the records outgrow their first allocation, so gates are run again after growth
    print("48-qubit ansatz adjoint", flush=True)
    wl = workloads.ansatz_line_48q()
    loss, grad = pp.adjointgradient(wl["circuit"], wl["obs"], wl["thetas"], max_weight=wl["max_weight"],
                                    min_abs_coeff=wl["min_abs_coeff"])
    rloss, rgrad = oracle_adjoint(wl["circuit"], wl["nqubits"], wl["thetas"], wl["obs"].terms,
                                  max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"])
    assert loss == pytest.approx(rloss[0], rel=1e-12)
    scale = float(np.abs(rgrad).max())
    np.testing.assert_allclose(grad, rgrad[:, 0], rtol=0, atol=1e-11 * scale)


def test_adjoint_rejects_bad_arguments():
    n = 6
    circ, obs = _ansatz(n, 1, False)
    P = pp.countparameters(circ)
    with pytest.raises(ValueError):
        pp.adjointgradient(circ, obs, np.zeros(P + 1))
    with pytest.raises(NotImplementedError):
        pp.rewindgradient(circ, obs, np.zeros(P), overlapfunc=lambda s: 0.0)


@pytest.mark.parametrize("B,noise,checkpoint", [(1, False, 0), (1, False, 40), (5, True, 0), (5, True, 37)])
def test_adjoint_chained_and_unchained_bitwise(B, noise, checkpoint):
    # the recorded forward runs list each rotation's touched terms from the
    # merge before it and leave out the rotations that touch nothing (the
    # default), or branch and launch every rotation: the records, and so the
    # losses and gradients, are the same bit for bit
    print(f"adjoint chained vs unchained B={B} noise={noise} checkpoint={checkpoint}", flush=True)
    from pauliprop_hip.hipsum import HIPPauliSum, PackedCircuit
    n = 10
    circ, obs = _ansatz(n, 4, noise)
    P = pp.countparameters(circ)
    th = np.random.default_rng(11 + B).normal(0, 0.3, size=(P, B))
    pc = PackedCircuit(circ, th if B > 1 else th[:, 0], n, batch=B)
    k0, c0 = obs.to_arrays(pc.W, B)
    out = []
    for chain in (True, False):
        s = HIPPauliSum(n, batch=B, capacity=1024)
        s.tune(chain_tails=chain, skip_identity=chain)
        s.upload(k0, c0)
        out.append(pc.adjoint(s, max_weight=6, min_abs_coeff=1e-6, checkpoint_gates=checkpoint))
    (l1, g1), (l2, g2) = out
    assert np.array_equal(l1.view(np.uint64), l2.view(np.uint64))
    assert np.array_equal(g1.view(np.uint64), g2.view(np.uint64))
    rloss, rgrad = oracle_adjoint(circ, n, th if B > 1 else th[:, 0], obs.terms, max_weight=6, min_abs_coeff=1e-6,
                                  batch=B)
    scale = max(1.0, float(np.abs(rgrad).max()))
    # Heisenberg parameter rows run in reverse Schrodinger order
    np.testing.assert_allclose(g1[::-1].reshape(rgrad.shape), rgrad, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("B", [1, 5])
def test_adjoint_independent_of_grid_per_cu(B):
    # the sweep's per-block partial gradients use a fixed block count, so the
    # gradients are the same bits whatever persistent grid the table chose
    print(f"adjoint over grid_per_cu, B={B}", flush=True)
    from pauliprop_hip.hipsum import HIPPauliSum, PackedCircuit
    n = 10
    circ, obs = _ansatz(n, 3, True)
    P = pp.countparameters(circ)
    th = np.random.default_rng(5 + B).normal(0, 0.3, size=(P, B))
    pc = PackedCircuit(circ, th if B > 1 else th[:, 0], n, batch=B)
    k0, c0 = obs.to_arrays(pc.W, B)
    out = []
    for g in (1, 4, 8):
        s = HIPPauliSum(n, batch=B, capacity=1024)
        s.tune(grid_per_cu=g)
        s.upload(k0, c0)
        out.append(pc.adjoint(s, max_weight=6, min_abs_coeff=1e-6))
    for loss, grad in out[1:]:
        assert np.array_equal(loss.view(np.uint64), out[0][0].view(np.uint64))
        assert np.array_equal(grad.view(np.uint64), out[0][1].view(np.uint64))
