"""Host-side logic without a GPU: topologies and builders, gate packing, the
C-ABI library's exports and struct layout."""
import ctypes
import os
import re
import subprocess
from collections import deque

import numpy as np
import pytest

import pauliprop_hip as pp
from pauliprop_hip import _lib
from pauliprop_hip.gates import pp_gate

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_heavyhex_builder_is_eagle():
    # src/Circuits/topologies.jl:185-200
    assert pp.heavyhextopology() == pp.ibmeagletopology
    edges = pp.heavyhextopology()
    nodes = {q for e in edges for q in e}
    assert len(edges) == 144 and nodes == set(range(1, 128))
    deg = {q: 0 for q in nodes}
    for a, b in edges:
        deg[a] += 1
        deg[b] += 1
    assert max(deg.values()) == 3


def _eccentricity(edges, src):
    adj = {}
    for a, b in edges:
        adj.setdefault(a, []).append(b)
        adj.setdefault(b, []).append(a)
    dist = {src: 0}
    dq = deque([src])
    while dq:
        u = dq.popleft()
        for v in adj[u]:
            if v not in dist:
                dist[v] = dist[u] + 1
                dq.append(v)
    return max(dist.values())


def test_heavyhex_observable_qubit_is_a_central_degree3_node():
    edges = pp.ibmeagletopology
    deg = {}
    for a, b in edges:
        deg[a] = deg.get(a, 0) + 1
        deg[b] = deg.get(b, 0) + 1
    d3 = [q for q, d in deg.items() if d == 3]
    ecc = {q: _eccentricity(edges, q) for q in d3}
    best = min(ecc.values())
    assert sorted(q for q in d3 if ecc[q] == best) == [63, 65] and deg[63] == 3


def test_topologies():
    b = pp.bricklayertopology(32, periodic=True)
    assert len(b) == 32 and b[15] == (31, 32) and (32, 1) in b and b[-1] == (32, 1)
    assert len(pp.rectangletopology(8, 8, periodic=True)) == 128
    assert len(pp.rectangletopology(6, 6)) == 60
    assert pp.staircasetopology(48) == [(i, i + 1) for i in range(1, 48)]


def test_circuit_sizes():
    # gate counts of the benchmark circuits
    c1 = pp.tfitrottercircuit(32, 32, topology=pp.bricklayertopology(32, periodic=True))
    assert len(c1) == 2048
    t = pp.rectangletopology(8, 8, periodic=True)
    c3 = []
    for _ in range(10):
        pp.rxxlayer(c3, t)
        pp.ryylayer(c3, t)
        pp.rzzlayer(c3, t)
        pp.rzlayer(c3, 64)
    assert len(c3) == 4480


def test_heisenberg_order_and_packing():
    circ = [pp.PauliRotation("X", 1), pp.DepolarizingNoise(2, 0.1), pp.PauliRotation(["Z", "Z"], [1, 2]),
            pp.CliffordGate("CNOT", [1, 2])]
    hc, hth = pp.toheisenberg(circ, [0.1, 0.2])
    assert isinstance(hc[0], pp.CliffordGate) and hth == [0.2, 0.1]
    gates, lc, ls, nl = pp.pack_circuit(hc, 2, 1)
    assert gates[0].kind == 3 and gates[1].kind == 0 and gates[1].param == 0 and gates[1].mask[0] == 0b1111
    assert gates[2].kind == 1 and gates[2].param == -1 and gates[2].value == 0.1 and gates[2].damp_codes == 0b1110
    assert gates[3].param == 1 and gates[3].mask[0] == 0b01
    assert nl == 1 and len(ls) == 16


def test_argument_checks():
    with pytest.raises(ValueError):
        pp.PauliRotation("X", 5).mask(4)
    with pytest.raises(ValueError):
        pp.DepolarizingNoise(1, 1.5)
    with pytest.raises(ValueError):
        pp.CliffordGate("CNOT", [1])


def _header_functions():
    text = open(os.path.join(ROOT, "include", "pphip.h")).read()
    return re.findall(r"^\s*(?:int|const char \*)\s*(pp_\w+)\s*\(", text, flags=re.M)


def test_library_exports_every_header_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpphip.so not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert {s[0] for s in _lib.SIGNATURES} == set(names)
    lib = _lib.lib()
    assert lib.pp_version() >= 100


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "pphip.h"\n#include "ppo.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(pp_gate),"
        " offsetof(pp_gate,kind), offsetof(pp_gate,param), offsetof(pp_gate,value), offsetof(pp_gate,qubit),"
        " offsetof(pp_gate,damp_codes), offsetof(pp_gate,nq), offsetof(pp_gate,qubits), offsetof(pp_gate,lut),"
        " sizeof(ppo_gate)); return 0;}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "oracle"), str(src), "-o",
                    str(exe)], check=True)
    vals = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(pp_gate), pp_gate.kind.offset, pp_gate.param.offset, pp_gate.value.offset,
            pp_gate.qubit.offset, pp_gate.damp_codes.offset, pp_gate.nq.offset, pp_gate.qubits.offset,
            pp_gate.lut.offset, ctypes.sizeof(pp_gate)]
    assert vals == want


def test_product_path_has_no_host_fallback():
    # without a HIP device the library refuses to create a sum; nothing runs on the CPU
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpphip.so not built")
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    with pytest.raises(_lib.PPError):
        pp.HIPPauliSum(8)


def test_tuning_table_and_autotuner_candidates():
    # the table pp_create reads: "<key_words> <max_batch> <name> <value>" lines
    # whose values the library accepts (pp_tune's ranges, include/pphip.h)
    import os
    from pauliprop_hip import _lib, autotune
    ok = {"merge_items": {0, 2, 3, 4}, "merge_defer": {0, 1}, "split_below": set(range(0, (1 << 31) + 1, 1 << 14)),
          "perm_items": {1, 2, 4},
          "batch_rows": {0, 16, 32, 64, 128, 256},
          "radix_bits": set(range(4, 9)), "grid_per_cu": set(range(1, 9)), "run_chunk": set(range(16, 1025))}
    assert set(ok) == set(_lib.TUNE_NAMES) == set(autotune.CANDIDATES) == set(autotune.DEFAULTS)
    for name, vals in autotune.CANDIDATES.items():
        assert set(vals) <= ok[name], name
        assert autotune.DEFAULTS[name] in ok[name]
    for B in (1, 32, 64, 256):
        assert all(r * B * 8 <= 64 * 1024 for r in autotune.candidates("batch_rows", B))
    classes = {(W, b) for W, b, _ in autotune.CLASSES}
    path = os.path.join(os.path.dirname(_lib.__file__), "pphip_tune.txt")
    if not os.path.exists(path):
        return
    seen = set()
    for line in open(path):
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        W, b, name, v = line.split()
        W, b, v = int(W), int(b), int(v)
        assert (W, b) in classes and name in ok and v in ok[name], line
        if name == "batch_rows":
            assert v * b * 8 <= 64 * 1024, line
        assert (W, b, name) not in seen, line
        seen.add((W, b, name))
