"""Autotuner of the launch shapes libpphip.so takes from its tuning table.

    python -m pauliprop_hip.autotune --out pauliprop_hip/pphip_tune.txt

(run from paulipropagation.jl_amd/ with the library built beside this file).

For each class of handle (key width, largest batch) it runs a representative
propagation on this GPU, tries every candidate of each launch shape in turn
(coordinate descent from the compiled defaults, keeping the fastest), checks
that every candidate gives the same keys and coefficients bit for bit, and
writes the table that pp_create reads (pp_load_tuning, include/pphip.h):

    <key_words> <max_batch> <name> <value>

Launch shapes: merge_items (the one-column merge's tile, 256 x 2..4
diagonals, with its partition), merge_defer (that merge storing each tile
one tile behind), split_below (sums, in terms, below which that merge searches
its own tile splits), perm_items (list entries per thread of the
tail permutation), batch_rows (term rows per tile of the batched merge),
radix_bits (digit width of the Clifford re-sort), grid_per_cu (workgroups per
compute unit of the persistent grids), run_chunk (gates enqueued between
overflow checks). The workgroup size is 256 throughout (four wave64), fixed
by the kernels' launch bounds.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import circuits as C, workloads
from .gates import CliffordGate, countparameters
from .hipsum import HIPPauliSum, PackedCircuit
from .paulis import PauliSum

CANDIDATES = {
    "merge_items": [2, 3, 4],
    "merge_defer": [0, 1],
    "split_below": [0, 1 << 14, 1 << 16, 1 << 18, 1 << 20, 1 << 22],
    "perm_items": [1, 2, 4],
    "grid_per_cu": [2, 4, 6, 8],
    "run_chunk": [64, 128, 256, 512],
    "batch_rows": [16, 32, 64, 128, 256],
    "radix_bits": [4, 5, 6, 7, 8],
}
# the library's compiled defaults (0: chosen by key width or batch)
DEFAULTS = {"merge_items": 0, "merge_defer": 0, "split_below": 1 << 16, "perm_items": 2, "batch_rows": 0, "radix_bits": 8, "grid_per_cu": 4, "run_chunk": 128}


class Case:
    """A propagation to time: a workload's packed circuit, observable and truncation."""

    def __init__(self, wl):
        self.name, self.nq, self.B = wl["name"], wl["nqubits"], wl.get("batch", 1)
        self.mw, self.mac = wl["max_weight"], wl["min_abs_coeff"]
        self.pc = PackedCircuit(wl["circuit"], wl["thetas"], self.nq, batch=self.B)
        self.k0, self.c0 = wl["obs"].to_arrays(self.pc.W, self.B)
        self.sum = HIPPauliSum(self.nq, batch=self.B, capacity=1 << 20)
        self.sum.tune(**DEFAULTS)  # start from the compiled defaults, whatever table is installed

    def run(self):
        self.sum.upload(self.k0, self.c0)
        self.pc.run(self.sum, max_weight=self.mw, min_abs_coeff=self.mac)
        self.sum.synchronize()

    def result(self):
        return self.sum.download()


def clifford_ansatz(nq: int, nlayers: int = 8):
    """RY and RZ on every qubit, then CNOT on a line, nlayers times: every CNOT
    re-sorts the sum (the radix sort's workload)."""
    circ = []
    for _ in range(nlayers):
        C.rylayer(circ, nq)
        C.rzlayer(circ, nq)
        for q in range(1, nq):
            circ.append(CliffordGate("CNOT", [q, q + 1]))
    th = workloads.SplitMix64(99).normals(countparameters(circ), 0.3)
    obs = PauliSum(nq)
    for i in range(1, nq):
        obs.add(["Z", "Z"], [i, i + 1], 1.0)
    return dict(name=f"clifford_ansatz_{nq}q", nqubits=nq, circuit=circ, thetas=th, obs=obs, max_weight=8,
                min_abs_coeff=1e-6)


def _first_step_b64():
    wl = workloads.heisenberg_square_64q_b64(nsteps=1)
    wl["name"] += "_first_step"
    return wl


# (key words, largest batch of the class, [(launch shape, workload)]); shapes a
# class does not use (the batched merge's rows at B = 1, the one-column merge's
# tile at B > 1) are left out
ONE_COLUMN = ("merge_items", "merge_defer", "split_below", "perm_items", "grid_per_cu", "run_chunk")
BATCHED = ("batch_rows", "perm_items", "grid_per_cu", "run_chunk")
CLASSES = [
    (4, 1, [(n, workloads.kicked_ising_heavyhex_127q) for n in ONE_COLUMN]),
    # split_below only acts on sums above 2^16 terms, which the 48-qubit ansatz
    # never reaches (67234 terms at most): the 6x6 TFIM grows to 1.04 M
    (2, 1, [(n, workloads.ansatz_line_48q) for n in ONE_COLUMN if n != "split_below"] +
     [("split_below", workloads.tfim_tilted_6x6_36q), ("radix_bits", lambda: clifford_ansatz(48))]),
    (1, 1, [(n, workloads.tfim_brick_32q) for n in ONE_COLUMN] + [("radix_bits", lambda: clifford_ansatz(24))]),
    (2, 32, [(n, workloads.ansatz_line_noisy_48q_b32) for n in BATCHED]),
    (2, 64, [(n, _first_step_b64) for n in BATCHED]),
]


def effective_default(name: str, W: int, B: int) -> int:
    """The value a handle uses when the table does not name the shape."""
    if name == "merge_items":
        return 3 if W == 4 else 4
    if name == "batch_rows":
        return 256 if B <= 16 else 64 if B <= 64 else 32 if B <= 128 else 16
    return DEFAULTS[name]


def candidates(name: str, B: int):
    vals = CANDIDATES[name]
    if name == "batch_rows":
        vals = [v for v in vals if v * B * 8 <= 64 * 1024]
    return vals


def median_time(case: Case, reps: int) -> float:
    case.run()
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        case.run()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def same_result(a, b) -> bool:
    (ka, ca), (kb, cb) = a, b
    return ka.shape == kb.shape and np.array_equal(ka, kb) and np.array_equal(ca.view(np.uint64), cb.view(np.uint64))


def tune_class(W: int, bmax: int, plan, reps: int = 5, margin: float = 0.01, log=print):
    """Coordinate descent over the class's launch shapes, each from the values
    chosen before it; a candidate replaces the default only when it is faster
    by more than `margin` (timing noise is about a percent). Returns {name:
    (value, seconds, {candidate: seconds})}. Raises if any candidate changes a
    result bit."""
    cases = {}
    chosen = {}
    out = {}
    for name, make in plan:
        if make not in cases:
            cases[make] = Case(make())
        case = cases[make]
        case.sum.tune(**chosen)
        case.run()
        ref = case.result()
        times = {}
        for v in candidates(name, case.B):
            case.sum.tune(**{name: v})
            t = median_time(case, reps)
            if not same_result(case.result(), ref):
                raise RuntimeError(f"{case.name}: {name} = {v} changed the result")
            times[v] = t
            log(f"W={W} B<={bmax} {case.name} {len(ref[0])} terms: {name}={v} {t * 1e3:.2f} ms")
        best = effective_default(name, W, case.B)
        fastest = min(times, key=times.get)
        if best not in times or times[fastest] < (1.0 - margin) * times[best]:
            best = fastest
        case.sum.tune(**{name: best})
        chosen[name] = best
        out[name] = (best, times[best], times)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", required=True, help="tuning table to write")
    ap.add_argument("--reps", type=int, default=5, help="timed runs per candidate (median)")
    ap.add_argument("--margin", type=float, default=0.01, help="least relative gain that replaces a default")
    ap.add_argument("--classes", default=None, help="comma-separated W:B classes to tune (default all)")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    want = None if args.classes is None else {tuple(int(x) for x in c.split(":")) for c in args.classes.split(",")}
    results = {}
    lines = ["# pphip tuning table: <key_words> <max_batch> <name> <value>",
             "# written by python -m pauliprop_hip.autotune; every value was checked to give identical results"]
    for W, bmax, plan in CLASSES:
        if want is not None and (W, bmax) not in want:
            continue
        res = tune_class(W, bmax, plan, reps=args.reps, margin=args.margin, log=log)
        results[f"W{W}_B{bmax}"] = {n: {"value": v, "ms": round(t * 1e3, 3),
                                        "candidates_ms": {str(k): round(x * 1e3, 3) for k, x in ts.items()}}
                                    for n, (v, t, ts) in res.items()}
        lines += [f"{W} {bmax} {n} {v}" for n, (v, _, _) in res.items()]
    tmp = args.out + ".tmp"
    with open(tmp, "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(tmp, args.out)
    print(json.dumps({"table": args.out, "classes": results}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
