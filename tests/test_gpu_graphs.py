"""GPU parity of chunks captured in hipGraphs (PP_TUNE_GRAPHS): a propagation
repeated on one handle replays the chunks captured the first time. Every run,
capture or replay, with other work on the device in between, must equal the
eager run and the oracle bit for bit."""
import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.hipsum import HIPPauliSum, PackedCircuit

pytestmark = pytest.mark.gpu


def _circuit(nq, nlayers, seed, noise=True):
    circ = []
    topo = pp.bricklayertopology(nq, periodic=True)
    for layer in range(nlayers):
        pp.rxlayer(circ, nq)
        pp.rzzlayer(circ, topo)
        if noise and layer % 2 == 1:
            pp.depolarizinglayer(circ, nq, 0.02)
    th = np.random.default_rng(seed).uniform(0.1, 1.0, pp.countparameters(circ))
    return circ, th


def _obs(nq):
    obs = pp.PauliSum(nq)
    for q in (3, nq // 2, nq - 2):
        obs.add(["Z"], [q])
    return obs


@pytest.mark.parametrize("nq,lds", [(40, False), (40, True), (100, False)])
def test_replays_match_eager_and_oracle(nq, lds):
    print(f"graphs nq={nq} lds={lds}", flush=True)
    circ, th = _circuit(nq, 6, nq)
    obs = _obs(nq)
    kw = dict(max_weight=7, min_abs_coeff=1e-5)
    pc = PackedCircuit(circ, th, nq)
    k0, c0 = obs.to_arrays(pc.W, 1)
    eager = HIPPauliSum(nq, capacity=1 << 18)
    eager.tune(lds_runs=lds)
    eager.upload(k0, c0)
    pc.run(eager, **kw)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, **kw)
    assert_bitwise(*eager.download(), *ref.get(), "eager vs oracle")
    g = HIPPauliSum(nq, capacity=1 << 18)
    g.tune(lds_runs=lds, graphs=True)
    other = HIPPauliSum(nq, capacity=1 << 12)
    for rep in range(4):
        g.profile_control(reset=True)
        g.upload(k0, c0)
        counts = pc.run(g, **kw).copy()
        prof = g.profile()
        print(f"replay {rep}: launches {prof['graph_launches']} captures {prof['graph_captures']}", flush=True)
        assert_bitwise(*g.download(), *eager.download(), f"graph run {rep}")
        assert prof["graph_launches"] > 0
        if rep > 0:
            assert prof["graph_captures"] == 0, "a repeated propagation captured again"
        # other device work between replays
        other.upload(k0, c0)
        pc.run(other, max_weight=4, min_abs_coeff=1e-3)
        assert abs(g.overlapwithzero() - eager.overlapwithzero()) == 0.0
    assert counts.sum() > 0


def test_new_angles_capture_new_graphs():
    nq = 40
    circ, th = _circuit(nq, 4, 5)
    obs = _obs(nq)
    kw = dict(max_weight=6, min_abs_coeff=1e-5)
    g = HIPPauliSum(nq, capacity=1 << 18)
    g.tune(graphs=True, lds_runs=False)
    for th2 in (th, th * 0.5, th):
        pc = PackedCircuit(circ, th2, nq)
        k0, c0 = obs.to_arrays(pc.W, 1)
        g.upload(k0, c0)
        pc.run(g, **kw)
        ref, _ = oracle_propagate(circ, nq, th2, obs.terms, **kw)
        assert_bitwise(*g.download(), *ref.get(), "graph run with its own angles")


def test_growth_drops_graphs():
    # a handle too small for the sum: the first run grows it (graphs captured
    # before the growth are dropped), later runs replay at the grown size
    nq = 40
    circ, th = _circuit(nq, 6, 11)
    obs = _obs(nq)
    kw = dict(max_weight=7, min_abs_coeff=1e-5)
    pc = PackedCircuit(circ, th, nq)
    k0, c0 = obs.to_arrays(pc.W, 1)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, **kw)
    g = HIPPauliSum(nq, capacity=1024)
    g.tune(graphs=True, lds_runs=False)
    for rep in range(3):
        g.upload(k0, c0)
        pc.run(g, **kw)
        assert_bitwise(*g.download(), *ref.get(), f"run {rep} on a growing handle")
