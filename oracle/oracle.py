"""ctypes wrapper of the CPU oracle libppo.so (oracle/ppo.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker and as the timed host baseline.
The product package never imports it.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libppo.so")

_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)
_P = ctypes.c_void_p
_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", HERE, "libppo.so"], check=True, capture_output=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = [
            ("ppo_create", _P, [ctypes.c_int, ctypes.c_int]),
            ("ppo_destroy", None, [_P]),
            ("ppo_set", ctypes.c_int, [_P, _U64P, _DP, ctypes.c_size_t]),
            ("ppo_load_merge", ctypes.c_int, [_P, _U64P, _DP, ctypes.c_size_t, ctypes.c_int, ctypes.c_double]),
            ("ppo_size", ctypes.c_size_t, [_P]),
            ("ppo_sorted", ctypes.c_int, [_P]),
            ("ppo_get", None, [_P, _U64P, _DP]),
            ("ppo_rotation", ctypes.c_long, [_P, _U64P, _DP, _DP, ctypes.c_int, ctypes.c_double]),
            ("ppo_imaginary_rotation", ctypes.c_long, [_P, _U64P, _DP, _DP, ctypes.c_int, ctypes.c_int,
                                                       ctypes.c_double]),
            ("ppo_pauli_noise", None, [_P, ctypes.c_int, ctypes.c_uint, _DP, ctypes.c_int, ctypes.c_double]),
            ("ppo_amplitude_damping", ctypes.c_long, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                                      ctypes.c_double]),
            ("ppo_clifford", None, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _U64P, _DP, ctypes.c_int,
                                    ctypes.c_double]),
            ("ppo_overlap_zero", None, [_P, _DP]),
            ("ppo_overlap_kind", None, [_P, ctypes.c_int, _U64P, _DP]),
            ("ppo_max_abs", None, [_P, _DP]),
            ("ppo_truncate", ctypes.c_double, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            ("ppo_resample", ctypes.c_double, [_P, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, ctypes.c_double,
                                               _DP, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
            ("ppo_get_coeff", None, [_P, _U64P, _DP]),
            ("ppo_scalar_product", ctypes.c_int, [_P, _P, _DP]),
            ("ppo_run", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t, _DP, ctypes.c_size_t, _U64P, _DP,
                                       ctypes.c_int, ctypes.c_double, ctypes.POINTER(ctypes.c_int64)]),
            ("ppo_adjoint", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t, _DP, ctypes.c_size_t, _U64P, _DP,
                                           ctypes.c_int, ctypes.c_double, ctypes.c_int, _U64P, _DP, _DP]),
            ("ppo_rewind", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t, _DP, ctypes.c_size_t, _U64P, _DP,
                                          ctypes.c_int, ctypes.c_double, ctypes.c_int, _U64P, _P, ctypes.c_double,
                                          _DP, _DP]),
            ("ppo_anticommutes", ctypes.c_int, [_U64P, _U64P, ctypes.c_int]),
            ("ppo_sign_exponent", ctypes.c_int, [_U64P, _U64P, ctypes.c_int]),
            ("ppo_weight", ctypes.c_int, [_U64P, ctypes.c_int]),
            ("ppo_contains_xy", ctypes.c_int, [_U64P, ctypes.c_int]),
            ("ppo_xor_sort_tail", ctypes.c_int, [_U64P, _DP, ctypes.c_size_t, ctypes.c_int, ctypes.c_int, _U64P]),
            ("ppo_multi_create", _P, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            ("ppo_multi_destroy", None, [_P]),
            ("ppo_multi_set", ctypes.c_int, [_P, _U64P, _DP, ctypes.c_size_t]),
            ("ppo_multi_size", ctypes.c_size_t, [_P]),
            ("ppo_multi_get", None, [_P, _U64P, _DP]),
            ("ppo_multi_run", ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_size_t, _DP, ctypes.c_size_t, ctypes.c_int,
                                             ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]),
        ]
        for name, res, args in sig:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _mw(max_weight) -> int:
    if max_weight is None or (isinstance(max_weight, float) and math.isinf(max_weight)):
        return -1
    return int(max_weight)


def u64p(a):
    return a.ctypes.data_as(_U64P)


def dp(a):
    return a.ctypes.data_as(_DP)


class OracleSum:
    """A VectorPauliSum restated on the CPU: sorted keys, n x B coefficients."""

    def __init__(self, W: int, B: int = 1):
        self.W, self.B = W, B
        self._s = lib().ppo_create(W, B)

    def __del__(self):
        if getattr(self, "_s", None):
            lib().ppo_destroy(self._s)
            self._s = None

    def set(self, keys, coeffs):
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, self.B)
        if lib().ppo_set(self._s, u64p(keys), dp(coeffs), keys.shape[0]) != 0:
            raise ValueError("duplicate key")
        return self

    def load_merge(self, keys, coeffs, max_weight=math.inf, min_abs=0.0):
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, self.B)
        lib().ppo_load_merge(self._s, u64p(keys), dp(coeffs), keys.shape[0], _mw(max_weight), float(min_abs))
        return self

    def __len__(self):
        return lib().ppo_size(self._s)

    def get(self):
        n = len(self)
        keys = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        coeffs = np.zeros((max(n, 1), self.B), dtype=np.float64)
        lib().ppo_get(self._s, u64p(keys), dp(coeffs))
        return keys[:n], coeffs[:n]

    def get_sorted(self):
        """get(), rows in ascending key order (a Clifford leaves the oracle's
        sum unsorted until the next branching gate, as the reference does)."""
        keys, coeffs = self.get()
        if len(keys) > 1:
            order = np.lexsort(tuple(keys[:, l] for l in range(keys.shape[1])))
            keys, coeffs = keys[order], coeffs[order]
        return keys, coeffs

    def clifford(self, qubits0, changes, signs, max_weight=math.inf, min_abs=0.0):
        q = (ctypes.c_int * 2)(*(list(qubits0) + [0] * (2 - len(qubits0))))
        c = np.ascontiguousarray(changes, dtype=np.uint64)
        sg = np.ascontiguousarray(signs, dtype=np.float64)
        lib().ppo_clifford(self._s, len(qubits0), q, u64p(c), dp(sg), _mw(max_weight), float(min_abs))

    def rotation(self, mask_limbs, cosv, sinv, max_weight=math.inf, min_abs=0.0):
        m = np.ascontiguousarray(mask_limbs, dtype=np.uint64)
        c = np.ascontiguousarray(np.broadcast_to(cosv, (self.B,)), dtype=np.float64)
        s = np.ascontiguousarray(np.broadcast_to(sinv, (self.B,)), dtype=np.float64)
        return lib().ppo_rotation(self._s, u64p(m), dp(c), dp(s), _mw(max_weight), float(min_abs))

    def imaginary_rotation(self, mask_limbs, coshv, sinhv, normalize=True, max_weight=math.inf, min_abs=0.0):
        m = np.zeros(4, dtype=np.uint64)
        m[:len(mask_limbs)] = mask_limbs
        c = np.ascontiguousarray(np.broadcast_to(coshv, (self.B,)), dtype=np.float64)
        s = np.ascontiguousarray(np.broadcast_to(sinhv, (self.B,)), dtype=np.float64)
        return lib().ppo_imaginary_rotation(self._s, u64p(m), dp(c), dp(s), int(bool(normalize)), _mw(max_weight),
                                            float(min_abs))

    def pauli_noise(self, qubit0, damp_codes, factor, max_weight=math.inf, min_abs=0.0):
        f = np.ascontiguousarray(np.broadcast_to(factor, (self.B,)), dtype=np.float64)
        lib().ppo_pauli_noise(self._s, int(qubit0), int(damp_codes), dp(f), _mw(max_weight), float(min_abs))

    def amplitude_damping(self, qubit0, gamma, max_weight=math.inf, min_abs=0.0):
        return lib().ppo_amplitude_damping(self._s, int(qubit0), float(gamma), _mw(max_weight), float(min_abs))

    def overlap_zero(self):
        out = np.zeros(self.B, dtype=np.float64)
        lib().ppo_overlap_zero(self._s, dp(out))
        return out

    def overlap_kind(self, kind: int, ones=None):
        """kind 0 |0..0>, 1 |+..+>, 2 computational with `ones` (W limbs, bit 2q per qubit in |1>)."""
        o = np.ascontiguousarray(ones if ones is not None else np.zeros(self.W), dtype=np.uint64)
        out = np.zeros(self.B, dtype=np.float64)
        lib().ppo_overlap_kind(self._s, int(kind), u64p(o), dp(out))
        return out

    def resample(self, method: int, target: int, squared=False, offset=0.0, draws=None, calibrate=True, rtol=0.01,
                 atol=1.0):
        """method 0 semideterministic systematic, 1 systematic, 2 multinomial (draws ascending in [0, 1));
        returns the comb spacing (weight per draw)."""
        d = np.ascontiguousarray(draws if draws is not None else np.zeros(1), dtype=np.float64)
        sp = lib().ppo_resample(self._s, int(method), int(target), int(bool(squared)), float(offset), dp(d),
                                int(bool(calibrate)), float(rtol), float(atol))
        if sp < 0:
            raise ValueError("bad resampling arguments")
        return sp

    def max_abs(self):
        out = np.zeros(self.B, dtype=np.float64)
        lib().ppo_max_abs(self._s, dp(out))
        return out

    def truncate(self, max_weight=math.inf, min_abs=0.0, min_rel=None):
        return lib().ppo_truncate(self._s, _mw(max_weight), float(min_abs), -1.0 if min_rel is None else float(min_rel))

    def get_coeff(self, key_limbs):
        k = np.ascontiguousarray(key_limbs, dtype=np.uint64)
        out = np.zeros(self.B, dtype=np.float64)
        lib().ppo_get_coeff(self._s, u64p(k), dp(out))
        return out

    def scalar_product(self, other: "OracleSum"):
        out = np.zeros(self.B, dtype=np.float64)
        if lib().ppo_scalar_product(self._s, other._s, dp(out)) != 0:
            raise ValueError("scalar product of sums with different key widths or batches")
        return out

    def run(self, gates, ngates, theta, nparams, lut_changes, lut_signs, max_weight, min_abs):
        th = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        if th.size == 0:
            th = np.zeros(1)
        lc = np.ascontiguousarray(lut_changes if len(lut_changes) else [0], dtype=np.uint64)
        ls = np.ascontiguousarray(lut_signs if len(lut_signs) else [0.0], dtype=np.float64)
        counts = np.zeros(max(ngates, 1), dtype=np.int64)
        rc = lib().ppo_run(self._s, ctypes.cast(gates, ctypes.c_void_p), ngates, dp(th), nparams, u64p(lc), dp(ls),
                           _mw(max_weight), float(min_abs), counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        if rc != 0:
            raise ValueError("oracle rejected a gate")
        return counts[:ngates]


def oracle_propagate(circuit, nqubits, thetas, init_terms: dict, max_weight=math.inf, min_abs_coeff=1e-10,
                     batch: int = 1, W: int | None = None, heisenberg: bool = True, normalize_coeffs: bool = True,
                     min_rel_coeff=None):
    """Propagate a circuit (given in Schrodinger order) on the oracle; returns
    (OracleSum, per-gate n_in counts). thetas: P or P x B. heisenberg=False
    propagates a state in circuit order with transposed gates."""
    import sys
    pkg = os.path.join(os.path.dirname(HERE), "paulipropagation.jl_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from pauliprop_hip.gates import countparameters, pack_circuit, toheisenberg, toschrodinger
    from pauliprop_hip.paulis import key_words, int_to_limbs

    W = W or key_words(nqubits)
    P = countparameters(circuit)
    th = np.zeros((0, batch)) if thetas is None or P == 0 else np.asarray(thetas, dtype=np.float64).reshape(P, -1)
    hc, hth = (toheisenberg if heisenberg else toschrodinger)(circuit, list(th))
    hth = np.array(hth).reshape(P, -1) if P else np.zeros((0, batch))
    gates, lc, ls, _ = pack_circuit(hc, nqubits, W, normalize_coeffs=normalize_coeffs)
    s = OracleSum(W, batch)
    keys = np.array([int_to_limbs(t, W) for t in init_terms], dtype=np.uint64).reshape(-1, W)
    coeffs = np.array([np.broadcast_to(c, (batch,)) for c in init_terms.values()], dtype=np.float64).reshape(-1, batch)
    s.set(keys, coeffs)
    if min_rel_coeff is None:
        counts = s.run(gates, len(hc), hth, P, lc, ls, max_weight, min_abs_coeff)
        return s, counts
    # a relative threshold: every gate but a Clifford branches and merges without
    # truncating, then truncate! runs with max(min_rel * maxabscoeff, min_abs)
    # (src/Propagation/specializations.jl:367-377, :225-229); a Clifford truncates
    # as without it (:154-177)
    counts = np.zeros(len(hc), dtype=np.int64)
    size = ctypes.sizeof(gates._type_)
    k = 0
    for g in range(len(hc)):
        one = ctypes.cast(ctypes.addressof(gates) + g * size, ctypes.POINTER(type(gates[0])))
        npar = 1 if gates[g].param >= 0 else 0
        row = hth[gates[g].param:gates[g].param + 1] if npar else np.zeros((1, batch))
        save = gates[g].param
        if npar:
            gates[g].param = 0
        if gates[g].kind == 3:
            counts[g] = s.run(one, 1, row, 1, lc, ls, max_weight, min_abs_coeff)[0]
        else:
            counts[g] = s.run(one, 1, row, 1, lc, ls, math.inf, 0.0)[0]
            s.truncate(max_weight, min_abs_coeff, min_rel_coeff)
        gates[g].param = save
    return s, counts


class MultiOracleSum:
    """The zone-parallel restatement of MultiPauliSum: 2^zbits zones, rotations
    and Pauli noise on nthreads OpenMP threads (the all-cores CPU baseline)."""

    def __init__(self, W: int, B: int = 1, zbits: int = 3):
        self.W, self.B = W, B
        self._s = lib().ppo_multi_create(W, B, zbits)
        if not self._s:
            raise ValueError("zbits must be in 0..8")

    def __del__(self):
        if getattr(self, "_s", None):
            lib().ppo_multi_destroy(self._s)
            self._s = None

    def set(self, keys, coeffs):
        keys = np.ascontiguousarray(keys, dtype=np.uint64).reshape(-1, self.W)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64).reshape(-1, self.B)
        if lib().ppo_multi_set(self._s, u64p(keys), dp(coeffs), keys.shape[0]) != 0:
            raise ValueError("duplicate Pauli string")
        return self

    def __len__(self):
        return lib().ppo_multi_size(self._s)

    def get_sorted(self):
        n = len(self)
        keys = np.zeros((max(n, 1), self.W), dtype=np.uint64)
        coeffs = np.zeros((max(n, 1), self.B), dtype=np.float64)
        lib().ppo_multi_get(self._s, u64p(keys), dp(coeffs))
        keys, coeffs = keys[:n], coeffs[:n]
        if n > 1:
            order = np.lexsort(tuple(keys[:, l] for l in range(self.W)))
            keys, coeffs = keys[order], coeffs[order]
        return keys, coeffs

    def run(self, gates_ptr, ngates, theta, nparams, max_weight, min_abs, nthreads, counts=None):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        cp = counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)) if counts is not None else None
        rc = lib().ppo_multi_run(self._s, gates_ptr, ngates, dp(th), nparams, _mw(max_weight), float(min_abs),
                                 int(nthreads), cp)
        if rc != 0:
            raise ValueError("the zone baseline runs rotations and Pauli noise only")


def oracle_adjoint(circuit, nqubits, thetas, init_terms: dict, max_weight=math.inf, min_abs_coeff=1e-10,
                   batch: int = 1, overlap: str = "zero", onebitinds=()):
    """(loss[B], grad[P x B]) of overlap(propagate(...)) w.r.t. the
    Schrodinger-ordered parameters, by the oracle's reverse sweep with every
    truncation held fixed; overlap "zero", "plus", "computational" (ones at
    `onebitinds`) or "maxmixed"."""
    import sys
    pkg = os.path.join(os.path.dirname(HERE), "paulipropagation.jl_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from pauliprop_hip.gates import countparameters, pack_circuit, toheisenberg
    from pauliprop_hip.paulis import key_words, int_to_limbs

    W = key_words(nqubits)
    P = countparameters(circuit)
    th = np.asarray(thetas, dtype=np.float64).reshape(P, -1)
    hc, hth = toheisenberg(circuit, list(th))
    hth = np.ascontiguousarray(np.array(hth).reshape(P, -1))
    gates, lc, ls, _ = pack_circuit(hc, nqubits, W)
    s = OracleSum(W, batch)
    keys = np.array([int_to_limbs(t, W) for t in init_terms], dtype=np.uint64).reshape(-1, W)
    coeffs = np.array([np.broadcast_to(c, (batch,)) for c in init_terms.values()], dtype=np.float64).reshape(-1, batch)
    s.set(keys, coeffs)
    loss = np.zeros(batch)
    grad = np.zeros((max(P, 1), batch))
    ones_int = 0
    for q in onebitinds:
        ones_int |= 1 << (2 * (int(q) - 1))
    ones = np.ascontiguousarray(int_to_limbs(ones_int, W), dtype=np.uint64)
    lcv = np.ascontiguousarray(lc if len(lc) else [0], dtype=np.uint64)
    lsv = np.ascontiguousarray(ls if len(ls) else [0.0], dtype=np.float64)
    rc = lib().ppo_adjoint(s._s, ctypes.cast(gates, ctypes.c_void_p), len(hc), dp(hth), P, u64p(lcv), dp(lsv),
                           _mw(max_weight), float(min_abs_coeff), READOUT_KINDS[overlap], u64p(ones), dp(loss),
                           dp(grad))
    if rc != 0:
        raise ValueError("oracle adjoint rejected a gate")
    # Heisenberg parameter rows run in reverse Schrodinger order
    return loss, grad[:P][::-1].copy()


READOUT_KINDS = {"zero": 0, "plus": 1, "computational": 2, "maxmixed": 3, "paulisum": 4}


def oracle_rewind(circuit, nqubits, thetas, init_terms: dict, max_weight=math.inf, min_abs_coeff=1e-10,
                  overlap: str = "zero", onebitinds=(), rho: dict | None = None):
    """(expectation, grad[P]) of rewindgradient(circuit, psum, params, overlapfunc;
    max_weight, min_abs_coeff) (src/gradient.jl:26-108), restated on the CPU.
    overlap: "zero", "plus", "computational" (ones at the 1-based `onebitinds`),
    "maxmixed", or "paulisum" (overlapwithpaulisum(rho, .), rho a dict of
    terms). The gradient is in Schrodinger parameter order, as the reference
    returns it."""
    import sys
    pkg = os.path.join(os.path.dirname(HERE), "paulipropagation.jl_amd")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from pauliprop_hip.gates import countparameters, pack_circuit, toheisenberg
    from pauliprop_hip.paulis import key_words, int_to_limbs

    W = key_words(nqubits)
    P = countparameters(circuit)
    th = np.asarray(thetas, dtype=np.float64).reshape(P)
    hc, hth = toheisenberg(circuit, list(th))
    hth = np.ascontiguousarray(np.array(hth, dtype=np.float64).reshape(-1) if P else np.zeros(1))
    gates, lc, ls, _ = pack_circuit(hc, nqubits, W)
    s = OracleSum(W, 1)
    keys = np.array([int_to_limbs(t, W) for t in init_terms], dtype=np.uint64).reshape(-1, W)
    s.set(keys, np.array(list(init_terms.values()), dtype=np.float64))
    kind = READOUT_KINDS[overlap]
    ones_int = 0
    for q in onebitinds:
        ones_int |= 1 << (2 * (int(q) - 1))
    ones = np.ascontiguousarray(int_to_limbs(ones_int, W), dtype=np.uint64)
    rs, scale = None, 1.0
    if kind == 4:
        rs = OracleSum(W, 1)
        rk = np.array([int_to_limbs(t, W) for t in rho], dtype=np.uint64).reshape(-1, W)
        rs.set(rk, np.array(list(rho.values()), dtype=np.float64))
        scale = 2.0 ** nqubits
    lcv = np.ascontiguousarray(lc if len(lc) else [0], dtype=np.uint64)
    lsv = np.ascontiguousarray(ls if len(ls) else [0.0], dtype=np.float64)
    loss = np.zeros(1)
    grad = np.zeros(max(P, 1))
    rc = lib().ppo_rewind(s._s, ctypes.cast(gates, ctypes.c_void_p), len(hc), dp(hth), P, u64p(lcv), dp(lsv),
                          _mw(max_weight), float(min_abs_coeff), kind, u64p(ones), rs._s if rs else None,
                          scale, dp(loss), dp(grad))
    if rc != 0:
        raise ValueError("rewindgradient takes rotations and Clifford gates only (no noise)")
    return float(loss[0]), grad[:P][::-1].copy()
