"""The benchmark workloads: circuits, parameters and observables this backend
is measured on, with a portable seeded generator so that every host and
device run reads identical parameter bits.

  tfim_brick_32q             32 qubits, periodic brick TFIM, 32 layers, all
                             angles 0.1, Z16, max_weight 6, min_abs_coeff 1e-4
                             (the reference README example)
  kicked_ising_heavyhex_127q 127 qubits, heavy-hex (IBM Eagle) kicked Ising:
                             20 layers of RX(theta_h) on every qubit then
                             RZZ(-pi/2) on every edge, theta_h ~ U[0, pi/2] per
                             gate (seed 127001), Z63, max_weight 12,
                             min_abs_coeff 1e-5
  heisenberg_square_64q_b64  64 qubits, 8x8 periodic lattice Heisenberg
                             Trotter steps (RXX, RYY, RZZ on every edge, then
                             RZ on every qubit), 10 steps, B = 64 columns
                             theta ~ U[0, 0.2] (seed 6401), observable
                             sum_edges Z_i Z_j, max_weight 8, min_abs_coeff 1e-6
  ansatz_line_48q            48 qubits, 12 layers of RY, RZ on every qubit then
                             RZZ on the open line, theta ~ N(0, 0.3^2) (seed
                             4801), sum_i Z_i Z_{i+1}, max_weight 7,
                             min_abs_coeff 1e-5
  ansatz_line_noisy_48q_b32  the same ansatz plus DepolarizingNoise(q, 0.01) on
                             every qubit after each layer, B = 32 columns
                             (seed 4802)
  tfim_tilted_6x6_36q        36 qubits, the reference's published CPU
                             benchmark (examples/advanced_performance.ipynb:
                             92-118): 14 layers of RX(0.1) and RZ(0.1) on every
                             qubit then RZZ(0.05) on rectangletopology(6, 6),
                             Z21 Z22, no weight truncation, min_abs_coeff 2^-20
"""
from __future__ import annotations

import math

import numpy as np

from . import circuits as C
from .gates import countparameters
from .paulis import PauliString, PauliSum

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): state += 0x9E3779B97F4A7C15, then
    two xor-shift-multiply rounds. uniform() = (x >> 11) * 2^-53; normal() by
    Box-Muller on two uniforms."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def uniforms(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        return np.array([lo + (hi - lo) * self.uniform() for _ in range(n)], dtype=np.float64)

    def normals(self, n: int, sigma: float = 1.0) -> np.ndarray:
        out = np.empty(n, dtype=np.float64)
        i = 0
        while i < n:
            u1 = self.uniform()
            u2 = self.uniform()
            if u1 <= 0.0:
                continue
            r = math.sqrt(-2.0 * math.log(u1))
            out[i] = sigma * r * math.cos(2.0 * math.pi * u2)
            if i + 1 < n:
                out[i + 1] = sigma * r * math.sin(2.0 * math.pi * u2)
            i += 2
        return out


def tfim_brick_32q():
    n = 32
    circ = C.tfitrottercircuit(n, 32, topology=C.bricklayertopology(n, periodic=True))
    th = np.full(countparameters(circ), 0.1)
    obs = PauliSum.from_string(PauliString(n, "Z", 16))
    return dict(name="tfim_brick_32q", nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=6, min_abs_coeff=1e-4)


def kicked_ising_heavyhex_127q(nlayers: int = 20):
    n = 127
    topo = C.heavyhextopology()
    circ = []
    for _ in range(nlayers):
        C.rxlayer(circ, n)
        C.rzzlayer(circ, topo)
    rng = SplitMix64(127001)
    th = []
    for _ in range(nlayers):
        th.extend(rng.uniforms(n, 0.0, math.pi / 2))
        th.extend([-math.pi / 2] * len(topo))
    obs = PauliSum.from_string(PauliString(n, "Z", 63))
    return dict(name="kicked_ising_heavyhex_127q", nqubits=n, circuit=circ, thetas=np.array(th), obs=obs, max_weight=12,
                min_abs_coeff=1e-5)


def heisenberg_square_64q_b64(nsteps: int = 10, B: int = 64):
    n = 64
    topo = C.rectangletopology(8, 8, periodic=True)
    circ = []
    for _ in range(nsteps):
        C.rxxlayer(circ, topo)
        C.ryylayer(circ, topo)
        C.rzzlayer(circ, topo)
        C.rzlayer(circ, n)
    P = countparameters(circ)
    th = SplitMix64(6401).uniforms(P * B, 0.0, 0.2).reshape(P, B)
    obs = PauliSum(n)
    for a, b in topo:
        obs.add(["Z", "Z"], [a, b], 1.0)
    return dict(name="heisenberg_square_64q_b64", nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=8, min_abs_coeff=1e-6,
                batch=B)


def _hea(n=48, nlayers=12, noise=False):
    circ = []
    topo = C.staircasetopology(n)
    for _ in range(nlayers):
        C.rylayer(circ, n)
        C.rzlayer(circ, n)
        C.rzzlayer(circ, topo)
        if noise:
            C.depolarizinglayer(circ, n, 0.01)
    obs = PauliSum(n)
    for i in range(1, n):
        obs.add(["Z", "Z"], [i, i + 1], 1.0)
    return circ, obs


def ansatz_line_48q():
    n = 48
    circ, obs = _hea(n, 12, noise=False)
    P = countparameters(circ)
    th = SplitMix64(4801).normals(P, 0.3)
    return dict(name="ansatz_line_48q", nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=7, min_abs_coeff=1e-5)


def ansatz_line_noisy_48q_b32(B: int = 32):
    n = 48
    circ, obs = _hea(n, 12, noise=True)
    P = countparameters(circ)
    th = SplitMix64(4802).normals(P * B, 0.3).reshape(P, B)
    return dict(name="ansatz_line_noisy_48q_b32", nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=7, min_abs_coeff=1e-5,
                batch=B)


def tfim_tilted_6x6_36q(nlayers: int = 14):
    n = 36
    layer = []
    C.rxlayer(layer, n)
    C.rzlayer(layer, n)
    C.rzzlayer(layer, C.rectangletopology(6, 6))
    th = np.full(len(layer), 0.1)  # 2 dt h with dt 0.05, h 1
    th[2 * n:] *= 0.5  # 2 dt J with J 0.5
    obs = PauliSum.from_string(PauliString(n, ["Z", "Z"], [21, 22]))
    return dict(name="tfim_tilted_6x6_36q", nqubits=n, circuit=layer * nlayers, thetas=np.tile(th, nlayers), obs=obs,
                max_weight=math.inf, min_abs_coeff=2.0 ** -20)


WORKLOADS = {f.__name__: f for f in (tfim_brick_32q, kicked_ising_heavyhex_127q, heisenberg_square_64q_b64,
                                     ansatz_line_48q, ansatz_line_noisy_48q_b32, tfim_tilted_6x6_36q)}
