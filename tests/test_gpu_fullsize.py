"""Parity at the headline size: the kicked Ising workload on the 127-qubit
heavy-hex graph (5420 gates, UInt256 keys, max_weight 12, min_abs_coeff 1e-5;
the sum peaks at 4.78 M terms), propagated whole on the device and by the
zone-parallel C restatement of MultiPauliSum on the host cores, compared bit
for bit (keys and every coefficient bit), with the rotations that touch no
term left out on the device and launched on the host."""
import ctypes
import os
import time

import numpy as np
import pytest

from helpers import assert_bitwise
import oracle as orc
import pauliprop_hip as pp
from pauliprop_hip import workloads
from pauliprop_hip.hipsum import PackedCircuit

pytestmark = pytest.mark.gpu


def test_kicked_ising_heavyhex_127q_whole_run_bitwise():
    print("kicked Ising 127q, whole run: device vs zone-parallel oracle", flush=True)
    wl = workloads.kicked_ising_heavyhex_127q()
    nq, mw, mac = wl["nqubits"], wl["max_weight"], wl["min_abs_coeff"]
    pc = PackedCircuit(wl["circuit"], wl["thetas"], nq)
    keys, coeffs = wl["obs"].to_arrays(pc.W, 1)
    dev = pp.HIPPauliSum(nq, capacity=1 << 20)
    dev.upload(keys, coeffs)
    pc.run(dev, max_weight=mw, min_abs_coeff=mac)
    dk, dc = dev.download()
    assert int(pc.counts[:pc.ngates].max()) == 4778420
    threads = min(16, os.cpu_count() or 1)
    zbits = max(0, (threads - 1).bit_length())
    multi = orc.MultiOracleSum(pc.W, 1, zbits).set(keys, coeffs)
    gsz = ctypes.sizeof(pc.gates._type_)
    th = np.ascontiguousarray(pc.theta.reshape(-1))
    t = time.perf_counter()
    multi.run(ctypes.cast(ctypes.addressof(pc.gates), ctypes.c_void_p), pc.ngates, th, pc.nparams, mw, mac, threads)
    print(f"oracle: {time.perf_counter() - t:.1f} s on {threads} threads, {len(multi)} terms", flush=True)
    assert_bitwise(dk, dc, *multi.get_sorted(), "kicked Ising 127q whole run")
    assert len(dk) == 1534655
