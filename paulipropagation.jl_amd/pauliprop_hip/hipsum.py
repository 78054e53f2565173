"""HIPPauliSum: a Pauli sum resident in MI355X memory, and the propagate()
entry points that run on it.

Mirrors the reference's user API (src/Propagation/generics.jl:54-95,
src/stateoverlap.jl:35): `propagate(circuit, psum, thetas; max_weight,
min_abs_coeff)` with the circuit in Schrodinger order, reversed here into
Heisenberg order (src/Gates/heisenbergschrodinger.jl:20-49), and
`overlapwithzero(psum)`. Every step runs in libpphip.so on the GPU.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .gates import CLIFFORD, countparameters, pack_circuit, toheisenberg, toschrodinger
from .paulis import PauliString, PauliSum, int_to_limbs, key_words, limbs_to_int

_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)


def _u64p(a):
    return a.ctypes.data_as(_U64P)


def _dp(a):
    return a.ctypes.data_as(_DP)


def _max_weight_arg(max_weight) -> int:
    if max_weight is None or (isinstance(max_weight, float) and math.isinf(max_weight)):
        return -1
    mw = int(max_weight)
    if mw < 0:
        raise ValueError("max_weight must be non-negative")
    return mw


class HIPPauliSum:
    """Device Pauli sum: sorted, duplicate-free keys and an n x B FP64
    coefficient matrix (B = batch columns sharing one key set)."""

    def __init__(self, nqubits: int, batch: int = 1, capacity: int = 1024, key_words_: int = 0):
        L = _lib.lib()
        self.nqubits = int(nqubits)
        self.batch = int(batch)
        h = ctypes.c_void_p()
        _lib.check(L.pp_create(self.nqubits, int(key_words_), self.batch, int(capacity), ctypes.byref(h)))
        self._h = h
        W = ctypes.c_int()
        L.pp_info(self._h, None, ctypes.byref(W), None, None, None)
        self.W = W.value

    # -- construction --------------------------------------------------------------
    @classmethod
    def from_paulisum(cls, psum, batch: int = 1, capacity: int = 1024) -> "HIPPauliSum":
        if isinstance(psum, PauliString):
            psum = PauliSum.from_string(psum)
        s = cls(psum.nqubits, batch=batch, capacity=max(capacity, 2 * len(psum)))
        keys, coeffs = psum.to_arrays(s.W, batch)
        s.upload(keys, coeffs)
        return s

    def upload(self, keys: np.ndarray, coeffs: np.ndarray) -> None:
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, self.batch)
        if keys.shape[0] != coeffs.shape[0]:
            raise ValueError("keys and coeffs have different lengths")
        _lib.check(_lib.lib().pp_upload(self._h, _u64p(keys), _dp(coeffs), keys.shape[0]))

    def load_merge(self, keys: np.ndarray, coeffs: np.ndarray, max_weight=math.inf, min_abs_coeff=0.0) -> None:
        """Replace the sum with terms in any order, equal strings summed in input
        order, then truncated -- all on the device."""
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, self.batch)
        if keys.shape[0] != coeffs.shape[0]:
            raise ValueError("keys and coeffs have different lengths")
        _lib.check(_lib.lib().pp_load_merge(self._h, _u64p(keys), _dp(coeffs), keys.shape[0],
                                            _max_weight_arg(max_weight), float(min_abs_coeff)))

    def download(self):
        """(keys n x W uint64 ascending, coeffs n x B float64)"""
        n = len(self)
        keys = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        coeffs = np.zeros((max(n, 1), self.batch), dtype=np.float64)
        got = ctypes.c_size_t()
        _lib.check(_lib.lib().pp_download(self._h, _u64p(keys), _dp(coeffs), max(n, 1), ctypes.byref(got)))
        return keys[:n], coeffs[:n]

    def clone(self) -> "HIPPauliSum":
        """An independent device copy (pp_clone): what deepcopy of the sum gives,
        as propagate and rewindgradient take it (src/Base/propagate.jl:3)."""
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().pp_clone(self._h, ctypes.byref(h)))
        c = HIPPauliSum.__new__(HIPPauliSum)
        c.nqubits, c.batch, c.W, c._h = self.nqubits, self.batch, self.W, h
        return c

    def __deepcopy__(self, memo):
        return self.clone()

    def to_paulisum(self) -> PauliSum:
        keys, coeffs = self.download()
        return PauliSum.from_arrays(self.nqubits, keys, coeffs)

    def __len__(self) -> int:
        n = ctypes.c_size_t()
        _lib.check(_lib.lib().pp_info(self._h, None, None, None, ctypes.byref(n), None))
        return n.value

    @property
    def capacity(self) -> int:
        c = ctypes.c_size_t()
        _lib.check(_lib.lib().pp_info(self._h, None, None, None, None, ctypes.byref(c)))
        return c.value

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().pp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- single gates ----------------------------------------------------------------
    def apply_rotation(self, mask: int, theta, max_weight=math.inf, min_abs_coeff=1e-10):
        """applymergetruncate!(PauliRotation, cache, theta) for an integer generator mask."""
        th = np.broadcast_to(np.asarray(theta, dtype=np.float64), (self.batch,))
        c = np.ascontiguousarray([math.cos(t) for t in th], dtype=np.float64)
        s = np.ascontiguousarray([math.sin(t) for t in th], dtype=np.float64)
        m = np.array([(mask >> (64 * l)) & ((1 << 64) - 1) for l in range(self.W)], dtype=np.uint64)
        st = _lib.pp_gate_stats()
        _lib.check(_lib.lib().pp_apply_rotation(self._h, _u64p(m), _dp(c), _dp(s), _max_weight_arg(max_weight),
                                                float(min_abs_coeff), ctypes.byref(st)))
        return st

    def apply_imaginary_rotation(self, mask: int, tau, normalize_coeffs: bool = True, max_weight=math.inf,
                                 min_abs_coeff=1e-10):
        """applymergetruncate!(ImaginaryPauliRotation, cache, tau; normalize_coeffs) for an
        integer generator mask (src/Propagation/specializations.jl:86-104)."""
        tu = np.broadcast_to(np.asarray(tau, dtype=np.float64), (self.batch,))
        c = np.ascontiguousarray([math.cosh(t) for t in tu], dtype=np.float64)
        s = np.ascontiguousarray([math.sinh(t) for t in tu], dtype=np.float64)
        m = np.array([(mask >> (64 * l)) & ((1 << 64) - 1) for l in range(self.W)], dtype=np.uint64)
        st = _lib.pp_gate_stats()
        _lib.check(_lib.lib().pp_apply_imaginary_rotation(self._h, _u64p(m), _dp(c), _dp(s), int(bool(normalize_coeffs)),
                                                          _max_weight_arg(max_weight), float(min_abs_coeff),
                                                          ctypes.byref(st)))
        return st

    def apply_pauli_noise(self, qubit: int, damp_codes: int, lam, max_weight=math.inf, min_abs_coeff=1e-10):
        lam = np.broadcast_to(np.asarray(lam, dtype=np.float64), (self.batch,))
        f = np.ascontiguousarray(1 - lam, dtype=np.float64)
        st = _lib.pp_gate_stats()
        _lib.check(_lib.lib().pp_apply_pauli_noise(self._h, int(qubit) - 1, int(damp_codes), _dp(f),
                                                   _max_weight_arg(max_weight), float(min_abs_coeff),
                                                   ctypes.byref(st)))
        return st

    def apply_amplitude_damping(self, qubit: int, gamma, max_weight=math.inf, min_abs_coeff=1e-10):
        """AmplitudeDampingNoise(qubit) with strength gamma (1-based qubit)."""
        g = np.ascontiguousarray(np.broadcast_to(np.asarray(gamma, dtype=np.float64), (self.batch,)))
        st = _lib.pp_gate_stats()
        _lib.check(_lib.lib().pp_apply_amplitude_damping(self._h, int(qubit) - 1, _dp(g),
                                                         _max_weight_arg(max_weight), float(min_abs_coeff),
                                                         ctypes.byref(st)))
        return st

    def apply_clifford(self, symbol: str, qinds, max_weight=math.inf, min_abs_coeff=1e-10):
        """CliffordGate(symbol, qinds) (1-based qubits); the sum is sorted again on the device."""
        from .gates import CliffordGate, clifford_table
        g = CliffordGate(symbol, qinds)
        qs = [q - 1 for q in g.qinds]
        if any(not 0 <= q < self.nqubits for q in qs):
            raise ValueError(f"Clifford qubits {g.qinds} out of range 1..{self.nqubits}")
        ch, sg = clifford_table(g.symbol, qs, self.W)
        c = np.ascontiguousarray(ch, dtype=np.uint64)
        s = np.ascontiguousarray(sg, dtype=np.float64)
        q = (ctypes.c_int * 2)(*(qs + [0] * (2 - len(qs))))
        st = _lib.pp_gate_stats()
        _lib.check(_lib.lib().pp_apply_clifford(self._h, len(qs), q, _u64p(c), _dp(s), _max_weight_arg(max_weight),
                                                float(min_abs_coeff), ctypes.byref(st)))
        return st

    def overlapwithzero(self):
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_overlap_zero(self._h, _dp(out)))
        return float(out[0]) if self.batch == 1 else out

    def _column(self, out):
        return float(out[0]) if self.batch == 1 else out

    def overlapwithplus(self):
        """<+|psum|+>: the coefficients of strings with no Y or Z (src/stateoverlap.jl:44)."""
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_overlap(self._h, _lib.PP_OVERLAP_PLUS, None, _dp(out)))
        return self._column(out)

    def overlapwithcomputational(self, onebitinds):
        """<x|psum|x> for the basis state with ones at the 1-based qubits
        `onebitinds` (src/stateoverlap.jl:59-80)."""
        ones = 0
        for q in onebitinds:
            q = int(q)
            if not 1 <= q <= self.nqubits:
                raise ValueError(f"qubit index {q} out of range 1..{self.nqubits}")
            ones |= 1 << (2 * (q - 1))
        m = np.ascontiguousarray(int_to_limbs(ones, self.W), dtype=np.uint64)
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_overlap(self._h, _lib.PP_OVERLAP_COMPUTATIONAL, _u64p(m), _dp(out)))
        return self._column(out)

    def getcoeff(self, pstr):
        """The coefficient(s) of a Pauli string (a PauliString or its integer
        term), zero when absent (src/Base/termsum.jl:104-124)."""
        term = pstr.term if isinstance(pstr, PauliString) else int(pstr)
        if term < 0 or term >> (2 * self.nqubits):
            raise ValueError("Pauli string acts beyond the sum's qubits")
        k = np.ascontiguousarray(int_to_limbs(term, self.W), dtype=np.uint64)
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_get_coeff(self._h, _u64p(k), _dp(out)))
        return self._column(out)

    def overlapwithmaxmixed(self):
        """Tr[psum I / 2^n]: the identity's coefficient (src/stateoverlap.jl:87-93)."""
        return self.getcoeff(0)

    def scalarproduct(self, other):
        """Sum over strings present in both of the coefficient products,
        Tr[psum1 psum2] / 2^n (src/stateoverlap.jl:120-135)."""
        if isinstance(other, PauliString):
            c = self.getcoeff(other)
            return c * other.coeff
        if not isinstance(other, HIPPauliSum):
            raise TypeError("scalarproduct of a HIPPauliSum needs a HIPPauliSum or a PauliString")
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_scalar_product(self._h, other._h, _dp(out)))
        return self._column(out)

    def maxabscoeff(self):
        """The largest |c| (per column) (src/Base/Primitives/maxabscoeff.jl:7)."""
        out = np.zeros(self.batch, dtype=np.float64)
        _lib.check(_lib.lib().pp_max_abs_coeff(self._h, _dp(out)))
        return self._column(out)

    def truncate(self, max_weight=math.inf, min_abs_coeff=1e-10, min_rel_coeff=None) -> float:
        """truncate!(psum; max_weight, min_abs_coeff, min_rel_coeff) on the device
        (src/Propagation/generics.jl:192-250); returns the coefficient threshold used."""
        thr = ctypes.c_double()
        rel = -1.0 if min_rel_coeff is None else float(min_rel_coeff)
        _lib.check(_lib.lib().pp_truncate(self._h, _max_weight_arg(max_weight), float(min_abs_coeff), rel,
                                          ctypes.byref(thr)))
        return thr.value

    def resample(self, target_size: int, resample_func=None, squared: bool = False, rng=None,
                 calibrate: bool = True, rtol: float = 0.01, atol: float = 1.0, offset=None, draws=None) -> float:
        """resample!(psum, target_size; resample_func, squared) on the device
        (src/Base/MonteCarlo/resample.jl:24-60). resample_func is one of
        semideterministic_systematic_resample (the default), systematic_resample,
        multinomial_resample (the default with squared=True). The random numbers
        -- the comb offset, or the multinomial draws -- come from `rng` (a numpy
        Generator) unless given. Returns the comb spacing (weight per draw)."""
        method = _resample_method(resample_func, squared)
        rng = np.random.default_rng() if rng is None else rng
        dp_draws = None
        if method == _lib.PP_RESAMPLE_MULTINOMIAL:
            d = np.sort(rng.random(int(target_size))) if draws is None else np.asarray(draws, dtype=np.float64)
            d = np.ascontiguousarray(d, dtype=np.float64)
            dp_draws = _dp(d)
            off = 0.0
        else:
            off = float(rng.random()) if offset is None else float(offset)
        sp = ctypes.c_double()
        _lib.check(_lib.lib().pp_resample(self._h, method, int(target_size), int(bool(squared)), off, dp_draws,
                                          int(bool(calibrate)), float(rtol), float(atol), ctypes.byref(sp)))
        return sp.value

    def synchronize(self) -> None:
        _lib.check(_lib.lib().pp_synchronize(self._h))

    def profile_control(self, timing: int = -1, reset: bool = False) -> None:
        """timing: 1 stamps the merge and noise kernels of run() on the device clock, 2 records HIP
        events around every kernel, 0 stops, -1 leaves it."""
        _lib.check(_lib.lib().pp_profile_control(self._h, int(timing), int(bool(reset))))

    def tune(self, spin_limit: int = None, dynamic_tiles: bool = None, chain_tails: bool = None,
             skip_identity: bool = None, adj_defer_bytes: int = None, force_rerun: bool = None,
             lds_runs: bool = None, graphs: bool = None, **shapes) -> None:
        """Bound of a look-back wait (polls); whether tiles are always claimed
        atomically; whether static-tile waits report running out (every gate
        then runs again with atomic claims: a test of that path); whether a
        merge lists the next rotation's touched terms; whether circuits run the
        rotations and noise of a small sum in shared memory, many gates per
        launch; whether chunks of gates are captured in hipGraphs and replayed
        when the same chunk comes again;
        whether circuits leave out the rotations that touch no term; the device
        memory the adjoint sweep may keep per-rotation partial gradients in.
        Launch shapes by name (merge_items, merge_defer, split_below, perm_items, batch_rows, radix_bits,
        grid_per_cu, run_chunk; any value the library accepts gives identical
        results)."""
        L = _lib.lib()
        for name, v in shapes.items():
            if name not in _lib.TUNE_NAMES:
                raise TypeError(f"unknown launch shape {name}")
            _lib.check(L.pp_tune(self._h, _lib.TUNE_NAMES[name], int(v)))
        if adj_defer_bytes is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_ADJ_DEFER_BYTES, int(adj_defer_bytes)))
        if skip_identity is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_SKIP_IDENTITY, int(bool(skip_identity))))
        if spin_limit is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_SPIN_LIMIT, int(spin_limit)))
        if dynamic_tiles is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_DYNAMIC_TILES, int(bool(dynamic_tiles))))
        if graphs is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_GRAPHS, int(bool(graphs))))
        if lds_runs is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_LDS_RUNS, int(bool(lds_runs))))
        if force_rerun is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_FORCE_RERUN, int(bool(force_rerun))))
        if chain_tails is not None:
            _lib.check(L.pp_tune(self._h, _lib.PP_TUNE_CHAIN_TAILS, int(bool(chain_tails))))

    def load_tuning(self, path: str) -> None:
        """Apply a tuning table (lines "<key_words> <max_batch> <name> <value>")."""
        _lib.check(_lib.lib().pp_load_tuning(self._h, path.encode()))

    def profile(self) -> dict:
        p = _lib.pp_profile()
        _lib.check(_lib.lib().pp_profile_read(self._h, ctypes.byref(p)))
        return {
            "gates": p.gates,
            "term_applications": p.term_applications,
            "fallbacks": p.fallbacks,
            "identity_gates": p.identity_gates,
            "persistent_gates": p.persistent_gates,
            "graph_launches": p.graph_launches,
            "graph_captures": p.graph_captures,
            "kernels": {k: {"launches": p.launches[i], "ms": p.milliseconds[i], "bytes": p.algorithmic_bytes[i]}
                        for i, k in enumerate(_lib.KERNEL_KINDS)},
        }


class PackedCircuit:
    """A Heisenberg-ordered circuit packed once into pp_gate descriptors and a
    P x B parameter matrix, for repeated runs (the benchmark's inner loop)."""

    def __init__(self, circuit, thetas, nqubits: int, batch: int = 1):
        P = countparameters(circuit)
        th = np.zeros((0, batch)) if thetas is None else np.asarray(thetas, dtype=np.float64).reshape(P, -1)
        hc, hth = toheisenberg(circuit, list(th))
        self.ngates = len(hc)
        self.nparams = P
        self.theta = np.zeros((max(P, 1), th.shape[1] if P else batch), dtype=np.float64)
        if P:
            self.theta[:P] = np.array(hth).reshape(P, -1)
        self.W = key_words(nqubits)
        self.gates, lc, ls, self.nluts = pack_circuit(hc, nqubits, self.W)
        self.lut_c = np.ascontiguousarray(lc or [0], dtype=np.uint64)
        self.lut_s = np.ascontiguousarray(ls or [0.0], dtype=np.float64)
        self.counts = np.zeros(max(self.ngates, 1), dtype=np.int64)

    def run(self, psum: HIPPauliSum, max_weight=math.inf, min_abs_coeff=1e-10, first: int = 0, last=None):
        """Run gates [first, last) on psum; self.counts[first:last] receives n_in per gate."""
        last = self.ngates if last is None else last
        g0 = ctypes.cast(ctypes.addressof(self.gates) + first * ctypes.sizeof(self.gates._type_),
                         ctypes.POINTER(type(self.gates[0])))
        c0 = self.counts[first:].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        _lib.check(_lib.lib().pp_run_circuit(
            psum._h, g0, last - first, _dp(self.theta), self.nparams, _u64p(self.lut_c), _dp(self.lut_s),
            self.nluts, _max_weight_arg(max_weight), float(min_abs_coeff), c0))
        return self.counts[first:last]


    def adjoint(self, psum: HIPPauliSum, max_weight=math.inf, min_abs_coeff=1e-10, checkpoint_gates: int = 0,
                overlapfunc=None, onebitinds=()):
        """(loss[B], grad[P x B]) of the overlap `overlapfunc` (default
        overlapwithzero) after the whole circuit, by the device adjoint sweep;
        grad rows are in Heisenberg parameter order. psum holds the observable
        and is left holding the propagated sum."""
        B = psum.batch
        if self.theta.shape[1] != B:
            raise ValueError(f"the circuit has {self.theta.shape[1]} parameter columns, the sum {B}")
        kind, ones = overlap_spec(overlapfunc, onebitinds, psum.nqubits, psum.W)
        loss = np.zeros(B)
        grad = np.zeros((max(self.nparams, 1), B))
        _lib.check(_lib.lib().pp_adjoint(
            psum._h, self.gates, self.ngates, _dp(self.theta), self.nparams, _u64p(self.lut_c), _dp(self.lut_s),
            self.nluts, _max_weight_arg(max_weight), float(min_abs_coeff), int(checkpoint_gates), kind, _u64p(ones),
            _dp(loss), _dp(grad)))
        return loss, grad[:self.nparams]


def run_heisenberg(psum: HIPPauliSum, hcircuit, hthetas, max_weight=math.inf, min_abs_coeff=1e-10,
                   counts: bool = False, normalize_coeffs: bool = True):
    """Run a circuit already in propagation order (Heisenberg, or Schrodinger
    from toschrodinger); hthetas is P x B (or P) in that order."""
    P = countparameters(hcircuit)
    B = psum.batch
    th = np.zeros((max(P, 1), B), dtype=np.float64)
    if P:
        th[:P] = np.asarray(hthetas, dtype=np.float64).reshape(P, -1)
    gates, lut_c, lut_s, nluts = pack_circuit(hcircuit, psum.nqubits, psum.W, normalize_coeffs=normalize_coeffs)
    lc = np.ascontiguousarray(lut_c or [0], dtype=np.uint64)
    ls = np.ascontiguousarray(lut_s or [0.0], dtype=np.float64)
    cnt = np.zeros(max(len(hcircuit), 1), dtype=np.int64)
    _lib.check(_lib.lib().pp_run_circuit(
        psum._h, gates, len(hcircuit), _dp(th), P, _u64p(lc), _dp(ls), nluts, _max_weight_arg(max_weight),
        float(min_abs_coeff), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
    return cnt[:len(hcircuit)] if counts else None


def propagate(circuit, psum, thetas=None, max_weight=math.inf, min_abs_coeff=1e-10, batch: int = None,
              counts: bool = False, heisenberg: bool = True, normalize_coeffs: bool = True, min_rel_coeff=None):
    """propagate(circuit, psum, thetas; max_weight, min_abs_coeff, heisenberg) on the GPU.

    `psum` is a PauliString, a PauliSum or a HIPPauliSum (propagated in place
    and returned). `thetas` holds one parameter per parametrized gate in
    Schrodinger order, or a P x B array for B batch columns. The reference's
    default min_abs_coeff is 1e-10 and max_weight Inf
    (src/Propagation/generics.jl:54). heisenberg=False propagates a state in
    the Schrodinger picture (circuit order, transposed gates,
    src/Propagation/generics.jl:78-86); normalize_coeffs applies to
    ImaginaryPauliRotation gates (src/Propagation/specializations.jl:86).
    min_rel_coeff (B = 1) truncates after every gate but a Clifford below
    max(min_rel_coeff * maxabscoeff, min_abs_coeff), gate by gate
    (src/Propagation/specializations.jl:367-377, :225-229).
    """
    P = countparameters(circuit)
    th = np.zeros((0, 1)) if thetas is None else np.asarray(thetas, dtype=np.float64)
    if th.ndim == 1:
        th = th.reshape(-1, 1)
    if th.shape[0] != P:
        raise ValueError(f"the circuit has {P} parameters, got {th.shape[0]}")
    B = batch or (th.shape[1] if P else 1)
    if not isinstance(psum, HIPPauliSum):
        psum = HIPPauliSum.from_paulisum(psum, batch=B)
    if P and th.shape[1] != psum.batch:
        raise ValueError(f"thetas have {th.shape[1]} columns, the sum has {psum.batch}")
    order = toheisenberg if heisenberg else toschrodinger
    hcirc, hth = order(circuit, list(th) if P else [])
    hth = np.array(hth) if P else None
    if min_rel_coeff is not None:
        cnt = _run_relative(psum, hcirc, hth, max_weight, min_abs_coeff, min_rel_coeff, normalize_coeffs)
        return (psum, cnt) if counts else psum
    cnt = run_heisenberg(psum, hcirc, hth, max_weight=max_weight, min_abs_coeff=min_abs_coeff, counts=counts,
                         normalize_coeffs=normalize_coeffs)
    return (psum, cnt) if counts else psum


def _run_relative(psum, hcirc, hth, max_weight, min_abs_coeff, min_rel_coeff, normalize_coeffs):
    """Gate by gate: a Clifford truncates as usual; any other gate applies and
    merges without truncation, then truncate! with the relative threshold."""
    if psum.batch != 1:
        raise ValueError("min_rel_coeff takes one coefficient column")
    P = countparameters(hcirc)
    th = np.zeros((max(P, 1), 1))
    if P:
        th[:P] = np.asarray(hth, dtype=np.float64).reshape(P, 1)
    gates, lut_c, lut_s, nluts = pack_circuit(hcirc, psum.nqubits, psum.W, normalize_coeffs=normalize_coeffs)
    lc = np.ascontiguousarray(lut_c or [0], dtype=np.uint64)
    ls = np.ascontiguousarray(lut_s or [0.0], dtype=np.float64)
    cnt = np.zeros(max(len(hcirc), 1), dtype=np.int64)
    size = ctypes.sizeof(gates._type_)
    for g in range(len(hcirc)):
        one = ctypes.cast(ctypes.addressof(gates) + g * size, ctypes.POINTER(type(gates[0])))
        clifford = gates[g].kind == CLIFFORD
        mw, mac = (max_weight, min_abs_coeff) if clifford else (math.inf, 0.0)
        _lib.check(_lib.lib().pp_run_circuit(psum._h, one, 1, _dp(th), P, _u64p(lc), _dp(ls), nluts,
                                             _max_weight_arg(mw), float(mac),
                                             cnt[g:].ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        if not clifford:
            psum.truncate(max_weight, min_abs_coeff, min_rel_coeff)
    return cnt[:len(hcirc)]


def adjointgradient(circuit, psum, thetas, max_weight=math.inf, min_abs_coeff=1e-10, checkpoint_gates: int = 0,
                    overlapfunc=None, onebitinds=()):
    """(expectation, gradient) of overlapfunc(propagate(circuit, psum, thetas))
    on the GPU by the adjoint sweep, with truncations held fixed as the forward
    run made them: the exact derivative of the truncated propagation, which
    central differences of propagate() measure. overlapfunc is overlapwithzero
    (the default), overlapwithplus, overlapwithmaxmixed or
    overlapwithcomputational with `onebitinds`. thetas is P (or P x B) in
    Schrodinger order and the gradient comes back in the same shape and order.
    Noise gates and Cliffords are allowed.

    This is not the reference's rewindgradient (src/gradient.jl:26-108), which
    truncates the operator again while it undoes the gates; both agree when
    nothing is truncated. `rewindgradient` below runs that algorithm."""
    P = countparameters(circuit)
    th = np.asarray(thetas, dtype=np.float64)
    vec = th.ndim == 1
    th = th.reshape(P, -1)
    B = th.shape[1]
    if not isinstance(psum, HIPPauliSum):
        psum = HIPPauliSum.from_paulisum(psum, batch=B)
    pc = PackedCircuit(circuit, th, psum.nqubits, batch=B)
    loss, grad = pc.adjoint(psum, max_weight=max_weight, min_abs_coeff=min_abs_coeff,
                            checkpoint_gates=checkpoint_gates, overlapfunc=overlapfunc, onebitinds=onebitinds)
    grad = grad[::-1].copy()  # Heisenberg parameter rows run in reverse Schrodinger order
    if vec:
        return float(loss[0]), grad[:, 0].copy()
    return loss, grad


def rewindgradient(circuit, psum, thetas, overlapfunc=None, max_weight=math.inf, min_abs_coeff=1e-10,
                   onebitinds=()):
    """rewindgradient(circuit, psum, params, overlapfunc; max_weight, min_abs_coeff)
    -> (expec, grad) (src/gradient.jl:26-108) on the GPU: the reference's own
    algorithm -- forward propagation, then the gates undone in their
    Schrodinger form on the truncated operator and on an untruncated dual sum,
    a commutator dot per parametrized rotation. overlapfunc is overlapwithzero
    (the default), overlapwithplus, overlapwithmaxmixed or
    overlapwithcomputational with `onebitinds`. psum is copied (the reference
    deep-copies it too); params in Schrodinger order, one column. Noise
    channels raise, as in the reference."""
    P = countparameters(circuit)
    th = np.asarray(thetas, dtype=np.float64).reshape(-1)
    if th.shape[0] != P:
        raise ValueError(f"the circuit has {P} parameters, got {th.shape[0]}")
    if isinstance(psum, HIPPauliSum):
        if psum.batch != 1:
            raise ValueError("rewindgradient takes one coefficient column")
        work = psum.clone()
    else:
        work = HIPPauliSum.from_paulisum(psum, batch=1)
    pc = PackedCircuit(circuit, th.reshape(P, 1), work.nqubits, batch=1)
    kind, ones = overlap_spec(overlapfunc, onebitinds, work.nqubits, work.W)
    loss = np.zeros(1)
    grad = np.zeros(max(P, 1))
    _lib.check(_lib.lib().pp_rewind(
        work._h, pc.gates, pc.ngates, _dp(pc.theta), pc.nparams, _u64p(pc.lut_c), _dp(pc.lut_s), pc.nluts,
        _max_weight_arg(max_weight), float(min_abs_coeff), kind, _u64p(ones), _dp(loss), _dp(grad)))
    return float(loss[0]), grad[:P][::-1].copy()


def overlap_spec(overlapfunc, onebitinds, nqubits: int, W: int):
    """(pp_overlap_kind, ones limbs) of an overlap function of this module."""
    ones_int = 0
    if overlapfunc is None or overlapfunc is overlapwithzero or overlapfunc == "zero":
        kind = _lib.PP_OVERLAP_ZERO
    elif overlapfunc is overlapwithplus or overlapfunc == "plus":
        kind = _lib.PP_OVERLAP_PLUS
    elif overlapfunc is overlapwithmaxmixed or overlapfunc == "maxmixed":
        kind = _lib.PP_OVERLAP_MAXMIXED
    elif overlapfunc is overlapwithcomputational or overlapfunc == "computational":
        kind = _lib.PP_OVERLAP_COMPUTATIONAL
        for q in onebitinds:
            q = int(q)
            if not 1 <= q <= nqubits:
                raise ValueError(f"qubit index {q} out of range 1..{nqubits}")
            ones_int |= 1 << (2 * (q - 1))
    else:
        raise NotImplementedError("the device gradients take overlapwithzero, overlapwithplus, "
                                  "overlapwithcomputational and overlapwithmaxmixed")
    return kind, np.ascontiguousarray(int_to_limbs(ones_int, W), dtype=np.uint64)


def overlapwithzero(psum):
    """<0|psum|0>: the sum of the coefficients of strings with no X or Y
    (src/stateoverlap.jl:12-35)."""
    if isinstance(psum, HIPPauliSum):
        return psum.overlapwithzero()
    raise TypeError("overlapwithzero runs on a HIPPauliSum; upload the sum first")


def semideterministic_systematic_resample(*args, **kwargs):
    """Name of the semideterministic systematic resampler (src/Base/MonteCarlo/resample.jl:139-170);
    pass it as resample_func."""
    raise TypeError("pass semideterministic_systematic_resample as resample_func of resample or mcpropagate")


def systematic_resample(*args, **kwargs):
    """Name of the systematic resampler (src/Base/MonteCarlo/resample.jl:96-112); pass it as resample_func."""
    raise TypeError("pass systematic_resample as resample_func of resample or mcpropagate")


def multinomial_resample(*args, **kwargs):
    """Name of the multinomial resampler (src/Base/MonteCarlo/resample.jl:64-78); pass it as resample_func."""
    raise TypeError("pass multinomial_resample as resample_func of resample or mcpropagate")


def _resample_method(resample_func, squared):
    if resample_func is None:
        return _lib.PP_RESAMPLE_MULTINOMIAL if squared else _lib.PP_RESAMPLE_SEMIDETERMINISTIC
    table = {semideterministic_systematic_resample: _lib.PP_RESAMPLE_SEMIDETERMINISTIC,
             systematic_resample: _lib.PP_RESAMPLE_SYSTEMATIC, multinomial_resample: _lib.PP_RESAMPLE_MULTINOMIAL}
    if resample_func not in table:
        raise ValueError(f"unknown resample_func {resample_func!r}")
    return table[resample_func]


def maxabscoeff(psum):
    """maxabscoeff(psum) (src/Base/Primitives/maxabscoeff.jl:7)."""
    return _needs_device(psum, "maxabscoeff").maxabscoeff()


def truncate(psum, max_weight=math.inf, min_abs_coeff=1e-10, min_rel_coeff=None):
    """truncate!(psum; max_weight, min_abs_coeff, min_rel_coeff) in place; returns the sum."""
    _needs_device(psum, "truncate").truncate(max_weight, min_abs_coeff, min_rel_coeff)
    return psum


def resample(psum, target_size, **kwargs):
    """resample!(psum, target_size; resample_func, squared) in place on a HIPPauliSum; returns it."""
    _needs_device(psum, "resample").resample(target_size, **kwargs)
    return psum


def mcpropagate(circuit, psum, thetas=None, max_size=None, resampling_size=None, resample_func=None,
                squared: bool = False, rng=None, max_weight=math.inf, min_abs_coeff=1e-10, heisenberg: bool = True,
                **resample_kwargs):
    """mcpropagate(circuit, psum, thetas; max_size, resampling_size, resample_func)
    (src/Base/MonteCarlo/mcpropagate.jl): propagate gate by gate and, whenever the
    sum grows past max_size, resample it to resampling_size (default
    round(max_size / 2)). Returns the HIPPauliSum."""
    if max_size is None:
        raise TypeError("mcpropagate needs max_size")
    resampling_size = round(max_size / 2) if resampling_size is None else int(resampling_size)
    P = countparameters(circuit)
    th = np.zeros((0, 1)) if thetas is None else np.asarray(thetas, dtype=np.float64)
    if th.ndim == 1:
        th = th.reshape(-1, 1)
    if th.shape[0] != P:
        raise ValueError(f"the circuit has {P} parameters, got {th.shape[0]}")
    if not isinstance(psum, HIPPauliSum):
        psum = HIPPauliSum.from_paulisum(psum, batch=1)
    if psum.batch != 1:
        raise ValueError("mcpropagate takes one coefficient column")
    order = toheisenberg if heisenberg else toschrodinger
    hcirc, hth = order(circuit, list(th) if P else [])
    rng = np.random.default_rng() if rng is None else rng
    pth = np.zeros((max(P, 1), 1))
    if P:
        pth[:P] = np.asarray(hth, dtype=np.float64).reshape(P, 1)
    gates, lut_c, lut_s, nluts = pack_circuit(hcirc, psum.nqubits, psum.W)
    lc = np.ascontiguousarray(lut_c or [0], dtype=np.uint64)
    ls = np.ascontiguousarray(lut_s or [0.0], dtype=np.float64)
    cnt = np.zeros(1, dtype=np.int64)
    size = ctypes.sizeof(gates._type_)
    for g in range(len(hcirc)):
        g0 = ctypes.cast(ctypes.addressof(gates) + g * size, ctypes.POINTER(type(gates[0])))
        _lib.check(_lib.lib().pp_run_circuit(psum._h, g0, 1, _dp(pth), P, _u64p(lc), _dp(ls), nluts,
                                             _max_weight_arg(max_weight), float(min_abs_coeff),
                                             cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
        if len(psum) > max_size:
            psum.resample(resampling_size, resample_func=resample_func, squared=squared, rng=rng, **resample_kwargs)
    return psum


def overlapwithplus(psum):
    """<+|psum|+> (src/stateoverlap.jl:44)."""
    return _needs_device(psum, "overlapwithplus").overlapwithplus()


def overlapwithcomputational(psum, onebitinds):
    """<x|psum|x> with ones at `onebitinds` (src/stateoverlap.jl:59-80)."""
    return _needs_device(psum, "overlapwithcomputational").overlapwithcomputational(onebitinds)


def overlapwithmaxmixed(psum):
    """Tr[psum I / 2^n] (src/stateoverlap.jl:87-93)."""
    return _needs_device(psum, "overlapwithmaxmixed").overlapwithmaxmixed()


def getcoeff(psum, pstr):
    """Coefficient of a Pauli string in the sum (src/Base/termsum.jl:104-124)."""
    return _needs_device(psum, "getcoeff").getcoeff(pstr)


def scalarproduct(a, b):
    """Tr[a b] / 2^n for HIPPauliSums (or one PauliString) (src/stateoverlap.jl:120-144)."""
    if isinstance(a, PauliString) and isinstance(b, HIPPauliSum):
        a, b = b, a
    return _needs_device(a, "scalarproduct").scalarproduct(b)


def overlapwithpaulisum(rho, psum):
    """Tr[rho psum] = 2^n scalarproduct(rho, psum) (src/stateoverlap.jl:102-104)."""
    return scalarproduct(rho, psum) * (2.0 ** _needs_device(rho, "overlapwithpaulisum").nqubits)


def _needs_device(psum, name):
    if isinstance(psum, HIPPauliSum):
        return psum
    raise TypeError(f"{name} runs on a HIPPauliSum; upload the sum first")


def key_to_int(row) -> int:
    return limbs_to_int(row)


__all__ = ["HIPPauliSum", "propagate", "mcpropagate", "resample", "truncate", "maxabscoeff", "semideterministic_systematic_resample",
           "systematic_resample", "multinomial_resample", "overlapwithzero", "overlapwithplus", "overlapwithcomputational",
           "overlapwithmaxmixed", "overlapwithpaulisum", "getcoeff", "scalarproduct", "run_heisenberg", "key_words"]
