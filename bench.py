#!/usr/bin/env python3
"""Benchmark: term-gate applications per second of propagate() on one MI355X.

Default workload (N=1): 127 qubits on the heavy-hex (IBM Eagle) graph, 256-bit
keys, FP64 coefficients; 20 layers of RX(theta_h) on every qubit then
RZZ(-pi/2) on all 144 edges (5420 gates), theta_h ~ U[0, pi/2] per gate from
SplitMix64 seed 127001; observable Z63; max_weight 12, min_abs_coeff 1e-5.

A step is one full propagation of that circuit from the observable, which is
uploaded before the timed region; the sum stays in HBM throughout. `value` is
sum over gates of the live terms entering each gate, divided by the wall time
of the K timed steps. Rotations that touch no term are the identity; the
library finds most of them from per-qubit Pauli presence and launches nothing
for them (they still count with their input size). detail reports how many,
and the time of one step with every rotation launched, whose result is
compared bit for bit.

roofline: the kernel with the most device time, its algorithmic bytes per
launch ((K + 8B)(n_in + n_out) for a merge: one read and one write of the live
set) over its average launch time. Every merge dispatch of the last timed step
stamps its first workgroup's start and its last workgroup's end on the
device's constant-rate clock (100 MHz s_memrealtime); start/stop HIP events on
each dispatch were measured to add the host-visible completion of every timed
dispatch (about 7 us a merge, and the step slows by as much). traffic is the
HBM bytes per launch from the committed rocprof counter summary.

cpu_baseline: the zone-parallel C restatement of the reference's MultiPauliSum
on all host cores (OpenMP), timed on the heaviest layer of the same run from
the identical device state, with the single-thread C oracle (VectorPauliSum)
beside it; the GPU is timed on the same layer from the same state.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "paulipropagation.jl_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="kicked_ising_heavyhex_127q",
                    choices=["tfim_brick_32q", "kicked_ising_heavyhex_127q", "heisenberg_square_64q_b64",
                             "ansatz_line_48q", "ansatz_line_noisy_48q_b32", "tfim_tilted_6x6_36q"])
    ap.add_argument("--layers", type=int, default=None,
                    help="heisenberg_square_64q_b64 only: its first LAYERS steps (the full ten outgrow one GPU)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the single-thread CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint-gradient measurement")
    ap.add_argument("--adjoint", default=None, choices=["ansatz_line_48q", "ansatz_line_noisy_48q_b32"],
                    help="measure the adjoint of this workload only (default both)")
    ap.add_argument("--size-log", default=None,
                    help="run the workload in chunks of 64 gates and write every gate's input size (and capacity) to "
                         "this file until it ends or a device allocation fails; no timing")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device three times (scales the HBM counters)")
    args = ap.parse_args()

    if args.copy_calibration:
        import torch
        n = 32 << 20
        x = torch.ones(n, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_per_copy": n * 8}))
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(json.dumps({"error": "this backend runs on one GPU; multi-GPU sharding is designed but not built",
                          "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import pauliprop_hip as pp
    from pauliprop_hip import workloads
    from pauliprop_hip.hipsum import PackedCircuit

    if args.layers is not None and args.workload != "heisenberg_square_64q_b64":
        ap.error("--layers applies to heisenberg_square_64q_b64 only")
    wl = (workloads.WORKLOADS[args.workload]() if args.layers is None
          else workloads.heisenberg_square_64q_b64(nsteps=args.layers))
    nq, B = wl["nqubits"], wl.get("batch", 1)
    mw, mac = wl["max_weight"], wl["min_abs_coeff"]
    pc = PackedCircuit(wl["circuit"], wl["thetas"], nq, batch=B)
    keys0, coeffs0 = wl["obs"].to_arrays(pc.W, B)
    psum = pp.HIPPauliSum(nq, batch=B, capacity=1 << 20)

    def step():
        psum.upload(keys0, coeffs0)
        pc.run(psum, max_weight=mw, min_abs_coeff=mac)
        psum.synchronize()

    if args.size_log:
        return size_log(args.size_log, pp, pc, psum, keys0, coeffs0, mw, mac, B, args.workload)

    for w in range(args.warmup):
        t = time.perf_counter()
        step()
        log(f"warmup {w}: {time.perf_counter() - t:.3f} s, {len(psum)} terms")
    psum.profile_control(timing=0, reset=True)
    psum.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step()
        log(f"step {k}: {time.perf_counter() - t0:.3f} s cumulative")
    psum.synchronize()
    elapsed = time.perf_counter() - t0
    prof_t = psum.profile()
    # one more step, after the timed region, in which every merge kernel stamps
    # its device-clock start and end (profile mode 1: the stamps and their
    # per-gate reduction would otherwise add to the timed steps)
    psum.profile_control(timing=1, reset=True)
    step()
    psum.synchronize()
    prof = psum.profile()
    psum.profile_control(timing=0)
    counts = pc.counts[:pc.ngates].copy()
    apps = int(prof_t["term_applications"]) * B
    value = apps / elapsed
    overlap = psum.overlapwithzero()
    final_terms = len(psum)

    # dominant kernel and its roofline
    kern = max(prof["kernels"].items(), key=lambda kv: kv[1]["ms"])
    kname, kst = kern
    avg_ms = kst["ms"] / max(kst["launches"], 1)
    avg_bytes = kst["bytes"] / max(kst["launches"], 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    traffic, traffic_src = measured_traffic(args.workload, kname)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": f"k_{kname}",
                "avg_launch_us": round(avg_ms * 1e3, 2), "timing": "device clock, first workgroup start to last workgroup end of each dispatch", "avg_algorithmic_bytes": round(avg_bytes),
                "launches_per_step": kst["launches"]}
    kernels = {k: {"launches": v["launches"], "ms_timed_step": round(v["ms"], 2),
                   "GBps": round(v["bytes"] / v["ms"] / 1e6, 1) if v["ms"] > 0 else None}
               for k, v in prof["kernels"].items() if v["launches"]}

    # Transparency about the rotations left out as touching no term: the same
    # step with every rotation launched, timed once after the timed region, and
    # its result compared bit for bit.
    identity = int(prof.get("identity_gates", 0))
    psum.tune(skip_identity=False)
    t = time.perf_counter()
    step()
    ms_all = (time.perf_counter() - t) * 1e3
    k_all, c_all = psum.download()
    psum.tune(skip_identity=True)
    step()
    k_one, c_one = psum.download()
    same_all = bool(np.array_equal(k_all, k_one) and np.array_equal(c_all.view(np.uint64), c_one.view(np.uint64)))
    log(f"every rotation launched: {ms_all:.1f} ms per step, result bit-identical: {same_all}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(pp, pc, psum, keys0, coeffs0, counts, wl, args.cpu_seconds)
        cpu["published_workload"] = published_anchor(pp, PackedCircuit)

    adj = None
    if not args.no_adjoint:
        adj = adjoint_detail(pp, PackedCircuit, [args.adjoint] if args.adjoint else None)

    L = 127 + 144 if args.workload == "kicked_ising_heavyhex_127q" else None
    out = {
        "metric": "Pauli-term gate applications/s + % HBM peak; adjoint grads/s vs plain FD",
        "value": round(value, 1),
        "unit": "term-gate applications/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload_text(args.workload, wl, args.layers), "nqubits": nq, "key_bits": 64 * pc.W, "batch": B,
                   "gates": pc.ngates, "max_weight": None if mw == math.inf else mw, "min_abs_coeff": mac},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {"term_applications_per_step": apps // args.steps, "peak_terms": int(counts.max()),
                   "final_terms": final_terms, "overlapwithzero": overlap if B == 1 else list(overlap)[:4],
                   "gates_touching_terms": kernels.get("merge", {}).get("launches"), "kernels": kernels,
                   "rotations_left_out_as_identity": identity,
                   "ms_per_step_every_rotation_launched": round(ms_all, 3),
                   "every_rotation_launched_bit_identical": same_all,
                   "layer_gates": L, "adjoint": adj},
    }
    print(json.dumps(out))
    return 0


def measured_traffic(workload, kname):
    """HBM bytes per launch of the dominant kernel from the committed rocprof
    counter summary of this workload (FETCH_SIZE and WRITE_SIZE runs, each
    scaled by what the counter reads for a plain copy of known size), or None
    when no summary exists for it."""
    import re
    path = os.path.join(ROOT, "profiles", f"k_{kname}1_{workload}_traffic.md")
    if not os.path.exists(path):
        return None, None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.]+)", open(path).read())
    if not m:
        return None, None
    return round(float(m.group(1))), os.path.relpath(path, ROOT)


def adjoint_detail(pp, PackedCircuit, names=None):
    """Gradients/s of the adjoint sweep against plain forward finite differences
    (2P forward runs per gradient, h = 1e-6) on the same GPU: the 48-qubit
    line ansatz (1716 parameters, B = 1) and the same ansatz with depolarizing
    noise (B = 32 gradients in one batched sweep; its FD baseline is 32
    separate B = 1 gradients). Forward runs are timed and multiplied out;
    the agreement is measured on a few parameters."""
    import numpy as np
    from pauliprop_hip import workloads
    out = {}
    for name in names or ("ansatz_line_48q", "ansatz_line_noisy_48q_b32"):
        wl = workloads.WORKLOADS[name]()
        nq, mw, mac = wl["nqubits"], wl["max_weight"], wl["min_abs_coeff"]
        th = np.asarray(wl["thetas"], dtype=np.float64)
        B = 1 if th.ndim == 1 else th.shape[1]
        P = pp.countparameters(wl["circuit"])
        pc = PackedCircuit(wl["circuit"], th, nq, batch=B)
        k0, c0 = wl["obs"].to_arrays(pc.W, B)
        s = pp.HIPPauliSum(nq, batch=B, capacity=1 << 18)

        def adjoint():
            s.upload(k0, c0)
            return pc.adjoint(s, max_weight=mw, min_abs_coeff=mac)

        adjoint()
        t = time.perf_counter()
        reps = 3
        for _ in range(reps):
            loss, grad = adjoint()
        t_adj = (time.perf_counter() - t) / reps
        # plain forward of one column (the FD baseline's unit of work)
        th1 = th if B == 1 else th[:, 0]
        pf = PackedCircuit(wl["circuit"], th1, nq)
        s1 = pp.HIPPauliSum(nq, capacity=1 << 18)
        k1, c1 = wl["obs"].to_arrays(pf.W, 1)

        def forward(pcx):
            s1.upload(k1, c1)
            pcx.run(s1, max_weight=mw, min_abs_coeff=mac)
            return s1.overlapwithzero()

        forward(pf)
        t = time.perf_counter()
        for _ in range(reps):
            forward(pf)
        t_fwd = (time.perf_counter() - t) / reps
        # agreement with central differences on a few parameters of column 0,
        # through forward runs of the same batch (so the batch-max truncation
        # sees the same columns)
        def loss0(thx):
            s.upload(k0, c0)
            PackedCircuit(wl["circuit"], thx, nq, batch=B).run(s, max_weight=mw, min_abs_coeff=mac)
            ov = s.overlapwithzero()
            return float(ov if B == 1 else ov[0])

        h = 1e-6
        diffs = []
        for k in np.linspace(0, P - 1, 4).astype(int):
            tp, tm = th.copy(), th.copy()
            if B == 1:
                tp[k] += h
                tm[k] -= h
            else:
                tp[k, 0] += h
                tm[k, 0] -= h
            fd = (loss0(tp) - loss0(tm)) / (2 * h)
            # Heisenberg parameter rows run in reverse Schrodinger order
            diffs.append(abs(fd - float(grad[P - 1 - k, 0])))
        adj_rate = B / t_adj
        fd_rate = 1.0 / (2 * P * t_fwd)
        out[name] = {"params": P, "batch": B, "adjoint_s": round(t_adj, 4), "forward_s": round(t_fwd, 4),
                     "grads_per_s": round(adj_rate, 2), "fd_grads_per_s": round(fd_rate, 4),
                     "adjoint_over_fd": round(adj_rate / fd_rate, 1), "loss0": float(loss[0]),
                     "max_abs_diff_vs_fd": float(max(diffs)), "fd_params_checked": len(diffs)}
        log(f"adjoint {name}: {t_adj * 1e3:.1f} ms per sweep of {B} gradient(s), forward {t_fwd * 1e3:.1f} ms, "
            f"x{adj_rate / fd_rate:.0f} over FD, max |adj - fd| {max(diffs):.2e}")
    return out


def size_log(path, pp, pc, psum, keys0, coeffs0, mw, mac, B, workload):
    """Every gate's input size, chunk by chunk, until the run ends or fails
    (e.g. a device allocation the sum outgrows); one line per gate
    (gate index, input terms, capacity after the chunk), then a JSON summary."""
    psum.upload(keys0, coeffs0)
    status, err, g = "complete", None, 0
    bytes_per_term = 8 * pc.W + 8 * B
    with open(path, "w") as f:
        f.write(f"# {workload}: gate n_in capacity (B = {B}, {bytes_per_term} B of key and coefficients per term)\n")
        while g < pc.ngates:
            last = min(pc.ngates, g + 64)
            try:
                cnt = pc.run(psum, max_weight=mw, min_abs_coeff=mac, first=g, last=last)
            except pp._lib.PPError as e:
                status, err = "stopped", str(e)
                f.write(f"# stopped in gates {g}..{last - 1}: {err}\n")
                break
            cap = psum.capacity
            for i, c in enumerate(cnt):
                f.write(f"{g + i} {int(c)} {cap}\n")
            f.flush()
            log(f"gates {g}..{last - 1}: {int(cnt[-1])} terms entering the last, capacity {cap}")
            g = last
    n_end = len(psum) if status == "complete" else None
    print(json.dumps({"size_log": path, "workload": workload, "status": status, "gates_done": g,
                      "gates": pc.ngates, "terms_at_end": n_end, "error": err, "capacity": psum.capacity,
                      "bytes_per_term_key_and_coefficients": bytes_per_term}))
    return 0


def published_anchor(pp, PackedCircuit):
    """The workload the reference publishes CPU numbers for
    (examples/advanced_performance.ipynb:92-118 of the reference): the RNG-free
    6x6 tilted-field Ising Trotter circuit (36 qubits, UInt128 keys; per layer
    RX and RZ on every qubit and RZZ on rectangletopology(6, 6); dt 0.05, h 1,
    J 0.5), 14 layers, Z21 Z22, min_abs_coeff 2^-20. Timed here on the GPU,
    on one host thread (the C oracle) and on all host threads (the zone-parallel
    C restatement), all bit-identical; rates under both counting conventions:
    exact (sum of every gate's input terms) and the notebook's (gates per layer
    times the terms at each layer's start, :457), beside its published
    218 M/s (one thread, :307-308) and 372-403 M/s (8 threads, :348, :391)."""
    import ctypes
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    from pauliprop_hip import workloads
    wl = workloads.tfim_tilted_6x6_36q()
    nq, nlay = wl["nqubits"], 14
    L = len(wl["circuit"]) // nlay
    pc = PackedCircuit(wl["circuit"], wl["thetas"], nq)
    k0, c0 = wl["obs"].to_arrays(pc.W, 1)
    mac = wl["min_abs_coeff"]
    dev = pp.HIPPauliSum(nq, capacity=1 << 21)

    def gpu_run():
        dev.upload(k0, c0)
        pc.run(dev, min_abs_coeff=mac)
        dev.synchronize()

    gpu_run()
    reps = 5
    t = time.perf_counter()
    for _ in range(reps):
        gpu_run()
    gpu_s = (time.perf_counter() - t) / reps
    counts = pc.counts[:pc.ngates].copy()
    exact = int(counts.sum())
    notebook = int(sum(L * counts[i * L] for i in range(nlay)))
    gk, gc = dev.download()
    ov = dev.overlapwithzero()
    th_flat = np.ascontiguousarray(pc.theta.reshape(-1))
    o = orc.OracleSum(pc.W, 1).set(k0, c0)
    t = time.perf_counter()
    o.run(pc.gates, pc.ngates, pc.theta, pc.nparams, [], [], None, mac)
    one_s = time.perf_counter() - t
    ok, oc = o.get()
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    zones = threads if threads & (threads - 1) == 0 else 2 * (1 << (threads - 1).bit_length())
    m = orc.MultiOracleSum(pc.W, 1, zones.bit_length() - 1).set(k0, c0)
    t = time.perf_counter()
    m.run(ctypes.cast(pc.gates, ctypes.c_void_p), pc.ngates, th_flat, pc.nparams, -1, mac, threads)
    all_s = time.perf_counter() - t
    mk, mc = m.get_sorted()
    same = bool(np.array_equal(gk, ok) and np.array_equal(gc.view(np.uint64), oc.view(np.uint64)) and
                np.array_equal(gk, mk) and np.array_equal(gc.view(np.uint64), mc.view(np.uint64)))

    def rates(sec):
        return {"seconds": round(sec, 4), "Mps_exact": round(exact / sec / 1e6, 1),
                "Mps_notebook": round(notebook / sec / 1e6, 1)}

    out = {"workload": "6x6 tilted-field Ising, 14 layers, Z21 Z22, min_abs_coeff 2^-20 (36 qubits, UInt128)",
           "terms": len(gk), "overlapwithzero": ov, "applications_exact": exact, "applications_notebook": notebook,
           "bitwise_equal_gpu_oracle_zones": same,
           "gpu": rates(gpu_s), "one_thread": rates(one_s), "all_threads": {**rates(all_s), "threads": threads},
           "cpu_model": cpu_model(),
           "published": {"one_thread_Mps": 218.0, "eight_threads_Mps": [372.0, 403.2],
                         "source": "examples/advanced_performance.ipynb:307-308,348,391"}}
    out["gpu_over_all_threads"] = round(all_s / gpu_s, 2)
    out["gpu_over_published_8_threads"] = round(notebook / gpu_s / 1e6 / 403.2, 2)
    out["one_thread_over_published"] = round(notebook / one_s / 1e6 / 218.0, 3)
    log(f"6x6 TFIM: {len(gk)} terms, <Z21Z22> {ov!r}; GPU {gpu_s * 1e3:.1f} ms, one thread {one_s:.2f} s, "
        f"{threads} threads {all_s:.2f} s; bit-identical {same}")
    return out


def workload_text(name, wl, layers=None) -> str:
    if layers is not None:
        return (f"Heisenberg on the 8x8 periodic square lattice (64 qubits), first {layers} of 10 steps, B=64, "
                "sum ZZ, UInt128 keys")
    return {
        "tfim_brick_32q": "TFIM on a 32-qubit periodic brick, 32 layers, theta 0.1, Z16, UInt64 keys",
        "kicked_ising_heavyhex_127q": "kicked Ising on the 127-qubit heavy-hex graph, 20 x (RX(U[0,pi/2]) on 127 qubits + RZZ(-pi/2) on 144 "
                   "edges), Z63, UInt256 keys",
        "heisenberg_square_64q_b64": "Heisenberg on the 8x8 periodic square lattice (64 qubits), 10 steps, B=64, sum ZZ, UInt128 keys",
        "ansatz_line_48q": "hardware-efficient ansatz on a 48-qubit line, 12 x (RY, RZ on every qubit, RZZ on the "
                           "line), theta N(0, 0.3^2), sum Z_i Z_i+1, UInt128 keys",
        "ansatz_line_noisy_48q_b32": "the 48-qubit line ansatz with DepolarizingNoise(0.01) on every qubit after "
                                     "each layer, B=32, theta N(0, 0.3^2), sum Z_i Z_i+1, UInt128 keys",
        "tfim_tilted_6x6_36q": "tilted-field Ising on the 6x6 open square lattice (36 qubits), 14 x (RX(0.1), RZ(0.1) on "
                               "every qubit, RZZ(0.05) on 60 edges), Z21 Z22, min_abs_coeff 2^-20, UInt128 keys",
    }[name]


def cpu_baseline(pp, pc, psum, keys0, coeffs0, counts, wl, budget_s):
    """CPU baselines on the heaviest layer of the run, from the identical state:
    the zone-parallel C restatement of MultiPauliSum on all host cores (the
    reported baseline) and the single-thread C oracle (VectorPauliSum), each
    checked bit for bit against the GPU's result of that layer."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctypes
    import oracle as orc

    L = 127 + 144 if wl["name"] == "kicked_ising_heavyhex_127q" else max(1, pc.ngates // 20)
    nl = pc.ngates // L
    loads = [int(counts[i * L:(i + 1) * L].sum()) for i in range(nl)]
    layer = int(np.argmax(loads))
    mw, mac = wl["max_weight"], wl["min_abs_coeff"]
    B = psum.batch
    # device state at the start of `layer`
    psum.upload(keys0, coeffs0)
    pc.run(psum, max_weight=mw, min_abs_coeff=mac, first=0, last=layer * L)
    keys, coeffs = psum.download()
    # GPU on that layer
    psum.synchronize()
    t = time.perf_counter()
    pc.run(psum, max_weight=mw, min_abs_coeff=mac, first=layer * L, last=(layer + 1) * L)
    psum.synchronize()
    gpu_s = time.perf_counter() - t
    gk, gc = psum.download()
    gpu_rate = loads[layer] * B / gpu_s
    th = np.ascontiguousarray(pc.theta.reshape(-1))
    gsz = ctypes.sizeof(pc.gates._type_)

    def gptr(g):
        return ctypes.cast(ctypes.addressof(pc.gates) + g * gsz, ctypes.c_void_p)

    def same(k, c):
        return bool(np.array_equal(k, gk) and np.array_equal(c.view(np.uint64), gc.view(np.uint64)))

    # all cores: zones as MultiPauliSum makes them, one zone per thread (src/Base/MultiSum/multisumstorage.jl:88-101)
    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    zones = threads if threads & (threads - 1) == 0 else 2 * (1 << (threads - 1).bit_length())
    multi = orc.MultiOracleSum(pc.W, B, zones.bit_length() - 1).set(keys, coeffs)
    t = time.perf_counter()
    multi.run(gptr(layer * L), L, th, pc.nparams, mw, mac, threads)
    multi_s = time.perf_counter() - t
    multi_rate = loads[layer] * B / multi_s
    mk, mc = multi.get_sorted()
    del multi
    # one core: the oracle, in chunks until the budget is spent
    o = orc.OracleSum(pc.W, B).set(keys, coeffs)
    done_apps = 0
    t = time.perf_counter()
    g = layer * L
    end = (layer + 1) * L
    while g < end:
        nxt = min(end, g + 16)
        cnt = np.zeros(nxt - g, dtype=np.int64)
        rc = orc.lib().ppo_run(o._s, gptr(g), nxt - g, orc.dp(th), pc.nparams, orc.u64p(pc.lut_c), orc.dp(pc.lut_s),
                               orc._mw(mw), float(mac), cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        if rc != 0:
            raise RuntimeError("oracle rejected a gate")
        done_apps += int(cnt.sum())
        g = nxt
        if time.perf_counter() - t > budget_s:
            break
    cpu_s = time.perf_counter() - t
    single_rate = done_apps * B / cpu_s
    single_same = same(*o.get()) if g == end else None
    return {"value": round(multi_rate, 1), "unit": "term-gate applications/s", "cores": threads, "kind": "port",
            "sample": f"layer {layer} of {nl} ({L} gates, {loads[layer]} term-gate applications) from the identical "
                      f"{len(keys)}-term state; zone-parallel C restatement of MultiPauliSum ({zones} zones, "
                      f"{threads} OpenMP threads) on {cpu_model()}",
            "bitwise_equal_after_layer": same(mk, mc),
            "gpu_same_sample": round(gpu_rate, 1), "gpu_over_cpu": round(gpu_rate / multi_rate, 1),
            "single_thread": {"value": round(single_rate, 1), "cores": 1,
                              "sample": f"{g - layer * L} of {L} gates of that layer ({done_apps} applications), "
                                        "single-thread C oracle (VectorPauliSum)",
                              "gpu_over_cpu": round(gpu_rate / single_rate, 1),
                              "bitwise_equal_after_layer": single_same}}


if __name__ == "__main__":
    sys.exit(main())
