"""GPU parity of ImaginaryPauliRotation, Schrodinger-picture propagation and
TGate: bit for bit against the oracle (single gates of every key width,
batched or not, normalized or not, truncated or not; whole Schrodinger
circuits with rotations, imaginary rotations, transposed Cliffords and noise),
and against the reference's known answers (test/test_imaginary.jl: the
two-term result, the closed-form thermal energies of a product-state X
Hamiltonian)."""
import math

import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import OracleSum, oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip.paulis import int_to_limbs, symboltoint

pytestmark = pytest.mark.gpu


def _rand_sum(rng, nq, n, with_identity=True):
    terms = set()
    while len(terms) < n:
        t = 0
        for q in rng.choice(nq, size=int(rng.integers(0, min(nq, 7) + 1)), replace=False):
            t |= int(rng.integers(1, 4)) << (2 * int(q))
        terms.add(t)
    if with_identity:
        terms.add(0)
    return sorted(terms)


@pytest.mark.parametrize("nq", [5, 40, 127])
def test_known_answers_on_device(nq):
    # test/test_imaginary.jl:1-57
    print(f"imaginary rotation known answers nq={nq}", flush=True)
    rng = np.random.default_rng(nq + 1)
    for support in (1, 2, 3, 4):
        symbols = [str(s) for s in rng.choice(["X", "Y", "Z"], support)]
        qinds = [int(q) for q in rng.choice(np.arange(1, nq + 1), support, replace=False)]
        P = symboltoint(nq, symbols, qinds)
        tau = float(rng.uniform())
        gate = pp.ImaginaryPauliRotation(symbols, qinds)
        dev = pp.propagate([gate], pp.PauliString(nq, "I", 1), [tau], min_abs_coeff=0, heisenberg=False)
        assert len(dev) == 2
        assert dev.getcoeff(0) == pytest.approx(1.0, rel=1e-15)
        assert dev.getcoeff(P) == pytest.approx(-math.sinh(tau) / math.cosh(tau), rel=1e-14)
        dev = pp.propagate([gate], pp.PauliString(nq, "I", 1), [tau], min_abs_coeff=0, heisenberg=False,
                           normalize_coeffs=False)
        assert dev.getcoeff(0) == math.cosh(tau) and dev.getcoeff(P) == -math.sinh(tau)
        with pytest.raises(RuntimeError):
            pp.propagate([gate], pp.PauliString(nq, "I", 1), [tau])


@pytest.mark.parametrize("nq,B", [(20, 1), (60, 1), (120, 1), (60, 3)])
@pytest.mark.parametrize("normalize", [True, False])
def test_single_imaginary_rotation_matches_oracle(nq, B, normalize):
    print(f"imaginary rotation nq={nq} B={B} normalize={normalize}", flush=True)
    rng = np.random.default_rng(nq * 3 + B + int(normalize))
    terms = _rand_sum(rng, nq, 20000)
    W = pp.key_words(nq)
    keys = np.array([int_to_limbs(t, W) for t in terms], dtype=np.uint64).reshape(-1, W)
    coeffs = rng.normal(size=(len(terms), B))
    coeffs[0] = 1.0 + rng.uniform(size=B)  # the identity
    for support, mw, mac in ((1, math.inf, 0.0), (2, 7, 1e-4), (3, math.inf, 1e-6)):
        qinds = [int(q) for q in rng.choice(np.arange(1, nq + 1), support, replace=False)]
        symbols = [str(s) for s in rng.choice(["X", "Y", "Z"], support)]
        mask = symboltoint(nq, symbols, qinds)
        taus = rng.uniform(0.05, 0.8, size=B)
        dev = pp.HIPPauliSum(nq, batch=B, capacity=4 * len(terms))
        dev.upload(keys, coeffs)
        st = dev.apply_imaginary_rotation(mask, taus if B > 1 else float(taus[0]), normalize_coeffs=normalize,
                                          max_weight=mw, min_abs_coeff=mac)
        ref = OracleSum(W, B).set(keys, coeffs)
        ch = np.array([math.cosh(t) for t in taus])
        sh = np.array([math.sinh(t) for t in taus])
        nc = ref.imaginary_rotation(int_to_limbs(mask, W), ch, sh, normalize, mw, mac)
        assert st.n_touched == nc
        assert_bitwise(*dev.download(), *ref.get(), f"imaginary rotation {symbols}{qinds} B={B}")


@pytest.mark.parametrize("B", [1, 2])
def test_schrodinger_circuit_matches_oracle(B):
    # rotations (angles negated), imaginary rotations, transposed Cliffords and
    # noise in circuit order
    print(f"Schrodinger circuit B={B}", flush=True)
    nq = 14
    circ = []
    for layer in range(3):
        pp.rxlayer(circ, nq)
        for q in (2, 7, 11):
            circ.append(pp.ImaginaryPauliRotation(["Z", "Z"], [q, q + 1]))
        circ.append(pp.ImaginaryPauliRotation("X", layer + 4))
        for q in range(2, nq, 4):
            circ.append(pp.CliffordGate("CNOT", [q, q + 1]))
        circ.append(pp.CliffordGate("S", [layer + 1]))
        pp.rzzlayer(circ, pp.bricklayertopology(nq))
        pp.depolarizinglayer(circ, nq, 0.01)
    P = pp.countparameters(circ)
    th = np.random.default_rng(40 + B).uniform(-0.6, 0.6, size=(P, B))
    rho = pp.PauliSum(nq, {0: 1.0, symboltoint(nq, ["Z"], [5]): 0.3, symboltoint(nq, ["X", "Y"], [9, 10]): -0.2})
    thb = th if B > 1 else th[:, 0]
    for mw, mac in ((math.inf, 1e-5), (6, 1e-4)):
        dev = pp.propagate(circ, rho, thb, max_weight=mw, min_abs_coeff=mac, heisenberg=False)
        ref, _ = oracle_propagate(circ, nq, thb, rho.terms, max_weight=mw, min_abs_coeff=mac, batch=B,
                                  heisenberg=False)
        assert_bitwise(*dev.download(), *ref.get_sorted(), f"Schrodinger circuit B={B} max_weight={mw}")


def test_imaginary_time_cooling_closed_form_on_device():
    # test/test_imaginary.jl:91-131: -sum h tanh(beta h) after every layer, to 1e-9
    print("imaginary time cooling", flush=True)
    nq = 4
    h = [0.8, -0.5, 1.2, -0.3]
    H = pp.PauliSum(nq)
    for i in range(nq):
        H.add(["X"], [i + 1], h[i])
    dH = pp.HIPPauliSum.from_paulisum(H)
    circ = [pp.ImaginaryPauliRotation("X", i + 1) for i in range(nq)]
    rho = pp.HIPPauliSum.from_paulisum(pp.PauliString(nq, "I", 1))
    energies = [pp.scalarproduct(dH, rho)]
    for layer in range(1, 11):
        pp.propagate(circ, rho, h, min_abs_coeff=0, heisenberg=False)
        energies.append(pp.scalarproduct(dH, rho))
        assert energies[-1] == pytest.approx(-sum(x * math.tanh(layer * x) for x in h), abs=1e-9)
    assert energies == sorted(energies, reverse=True)


def test_tgate_is_a_frozen_z_rotation():
    # src/Propagation/specializations.jl:293-308: TGate(q) acts as PauliRotation(:Z, q) at pi/4
    print("T gate", flush=True)
    nq = 30
    circ, circ_rz = [], []
    for q in range(1, nq + 1):
        circ += [pp.PauliRotation("X", q), pp.TGate(q)]
        circ_rz += [pp.PauliRotation("X", q), pp.FrozenGate(pp.PauliRotation("Z", q), math.pi / 4)]
    pp.rzzlayer(circ, pp.bricklayertopology(nq))
    pp.rzzlayer(circ_rz, pp.bricklayertopology(nq))
    th = np.random.default_rng(3).uniform(-1, 1, size=pp.countparameters(circ))
    obs = pp.PauliSum(nq, {symboltoint(nq, ["Y"], [7]): 1.0, symboltoint(nq, ["Z", "X"], [12, 13]): 0.5})
    a = pp.propagate(circ, obs, th, max_weight=6, min_abs_coeff=1e-6)
    b = pp.propagate(circ_rz, obs, th, max_weight=6, min_abs_coeff=1e-6)
    ref, _ = oracle_propagate(circ, nq, th, obs.terms, max_weight=6, min_abs_coeff=1e-6)
    assert_bitwise(*a.download(), *b.download(), "TGate vs frozen RZ(pi/4)")
    assert_bitwise(*a.download(), *ref.get(), "TGate vs oracle")
    with pytest.raises(NotImplementedError):
        pp.propagate([pp.TGate(1)], pp.PauliString(nq, "Z", 1), heisenberg=False)
