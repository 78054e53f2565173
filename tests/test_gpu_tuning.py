"""GPU: every launch shape of the tuning table gives the oracle's result bit
for bit, and pp_load_tuning applies, rejects and selects table lines."""
import numpy as np
import pytest

from helpers import assert_bitwise
from oracle import oracle_propagate
import pauliprop_hip as pp
from pauliprop_hip import _lib, autotune
from pauliprop_hip.hipsum import HIPPauliSum, PackedCircuit

pytestmark = pytest.mark.gpu


def _tfim32(nlayers=32):
    n = 32
    circ = pp.tfitrottercircuit(n, nlayers, topology=pp.bricklayertopology(n, periodic=True))
    th = np.full(pp.countparameters(circ), 0.1)
    obs = pp.PauliSum.from_string(pp.PauliString(n, "Z", 16))
    return dict(nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=8, min_abs_coeff=1e-8)


def _ansatz48(nlayers=8):
    n = 48
    circ = []
    for _ in range(nlayers):
        pp.rylayer(circ, n)
        pp.rzlayer(circ, n)
        pp.rzzlayer(circ, pp.staircasetopology(n))
    th = np.random.default_rng(48).normal(0, 0.3, pp.countparameters(circ))
    obs = pp.PauliSum(n)
    for i in range(1, n):
        obs.add(["Z", "Z"], [i, i + 1], 1.0)
    return dict(nqubits=n, circuit=circ, thetas=th, obs=obs, max_weight=7, min_abs_coeff=1e-6)


def _kicked127(nlayers=8):
    from pauliprop_hip import workloads
    return workloads.kicked_ising_heavyhex_127q(nlayers)


def _cliff(nq):
    return autotune.clifford_ansatz(nq, nlayers=4)


def _heis(B, nlayers=2, noise=False):
    nq = 9
    topo = pp.rectangletopology(3, 3, periodic=True)
    circ = []
    for _ in range(nlayers):
        pp.rxxlayer(circ, topo)
        pp.ryylayer(circ, topo)
        pp.rzzlayer(circ, topo)
        pp.rzlayer(circ, nq)
        if noise:
            pp.depolarizinglayer(circ, nq, 0.02)
    th = np.random.default_rng(B).uniform(0, 0.2, size=(pp.countparameters(circ), B))
    obs = pp.PauliSum(nq)
    for a, b in topo:
        obs.add(["Z", "Z"], [a, b], 1.0)
    return dict(nqubits=nq, circuit=circ, thetas=th, obs=obs, max_weight=5, min_abs_coeff=1e-7, batch=B)


def _device(wl, table=None, **shapes):
    B = wl.get("batch", 1)
    pc = PackedCircuit(wl["circuit"], wl["thetas"], wl["nqubits"], batch=B)
    s = HIPPauliSum(wl["nqubits"], batch=B, capacity=1024)
    if table is not None:
        s.load_tuning(str(table))
    s.tune(lds_runs=False)  # the shapes are the per-gate pipeline's
    s.tune(**shapes)
    k, c = wl["obs"].to_arrays(pc.W, B)
    s.upload(k, c)
    pc.run(s, max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"])
    return s.download()


_ORACLE = {}


def _oracle(name, wl):
    if name not in _ORACLE:
        ref, _ = oracle_propagate(wl["circuit"], wl["nqubits"], wl["thetas"], wl["obs"].terms,
                                  max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"],
                                  batch=wl.get("batch", 1))
        _ORACLE[name] = ref.get()
    return _ORACLE[name]


_WL = {"tfim32": _tfim32, "ansatz48": _ansatz48, "kicked127": _kicked127}


@pytest.mark.parametrize("name", list(_WL))
@pytest.mark.parametrize("shape,value", [("merge_items", v) for v in (2, 3, 4)]
                         + [("perm_items", v) for v in (1, 2, 4)]
                         + [("grid_per_cu", v) for v in (1, 2, 8)]
                         + [("run_chunk", v) for v in (16, 1024)])
def test_one_column_shapes_bitwise(name, shape, value):
    wl = _WL[name]()
    got = _device(wl, **{shape: value})
    assert_bitwise(*got, *_oracle(name, wl), f"{name} {shape}={value}")


@pytest.mark.parametrize("nq", [24, 48, 100])
@pytest.mark.parametrize("bits", [4, 5, 6, 7, 8])
def test_radix_bits_bitwise(nq, bits):
    wl = _cliff(nq)
    got = _device(wl, radix_bits=bits)
    assert_bitwise(*got, *_oracle(f"cliff{nq}", wl), f"Clifford ansatz {nq}q radix_bits={bits}")


@pytest.mark.parametrize("B,rows", [(5, 16), (5, 256), (32, 16), (32, 64), (32, 256), (64, 32), (64, 128)])
@pytest.mark.parametrize("noise", [False, True])
def test_batch_rows_bitwise(B, rows, noise):
    wl = _heis(B, noise=noise)
    got = _device(wl, batch_rows=rows, perm_items=1)
    assert_bitwise(*got, *_oracle(f"heis{B}{noise}", wl), f"B={B} rows={rows} noise={noise}")


@pytest.mark.parametrize("shape,value", [("merge_items", 1), ("merge_items", 5), ("merge_defer", 2), ("split_below", -1),
                                         ("split_below", (1 << 31) + 1), ("perm_items", 3),
                                         ("batch_rows", 8), ("batch_rows", 48), ("batch_rows", 512),
                                         ("radix_bits", 3), ("radix_bits", 9), ("grid_per_cu", 0),
                                         ("grid_per_cu", 9), ("run_chunk", 8), ("run_chunk", 2048)])
def test_out_of_range_shapes_rejected(shape, value):
    s = HIPPauliSum(9, batch=1)
    with pytest.raises(ValueError):
        s.tune(**{shape: value})


def test_batch_rows_beyond_shared_memory_rejected():
    s = HIPPauliSum(9, batch=512)
    with pytest.raises(ValueError):
        s.tune(batch_rows=32)  # 32 x 512 x 8 bytes > 64 KiB
    s.tune(batch_rows=16)


def test_table_applies_to_its_class(tmp_path):
    wl = _tfim32()
    t = tmp_path / "t.txt"
    t.write_text("# comment\n\n1 1 merge_items 2\n1 1 perm_items 1\n1 1 run_chunk 16\n"
                 "2 1 merge_items 9\n"  # another class: never applied, so never validated
                 "1 8 radix_bits 3\n")  # a wider batch class: the B = 1 handle takes 1 1
    assert_bitwise(*_device(wl, table=t), *_oracle("tfim32", wl), "table")


def test_table_picks_smallest_covering_batch(tmp_path):
    t = tmp_path / "t.txt"
    t.write_text("1 4 batch_rows 512\n1 64 batch_rows 16\n")
    with pytest.raises(ValueError):  # B = 3 takes the "1 4" class, whose value is invalid
        HIPPauliSum(9, batch=3).load_tuning(str(t))
    HIPPauliSum(9, batch=5).load_tuning(str(t))  # B = 5 takes "1 64"
    HIPPauliSum(9, batch=65).load_tuning(str(t))  # no class covers 65: nothing applied


@pytest.mark.parametrize("text", ["1 1 merge_items\n", "one 1 merge_items 2\n", "1 1 tile_width 2\n"])
def test_malformed_table_rejected(tmp_path, text):
    t = tmp_path / "t.txt"
    t.write_text(text)
    with pytest.raises(ValueError):
        HIPPauliSum(9).load_tuning(str(t))


def test_missing_table_rejected(tmp_path):
    with pytest.raises(ValueError):
        HIPPauliSum(9).load_tuning(str(tmp_path / "none.txt"))


def test_environment_table_read_at_create(tmp_path, monkeypatch):
    t = tmp_path / "t.txt"
    t.write_text("1 1 radix_bits 12\n")
    monkeypatch.setenv("PPHIP_TUNE_FILE", str(t))
    with pytest.raises(ValueError):
        HIPPauliSum(9)
    t.write_text("1 1 radix_bits 5\n1 1 merge_items 2\n")
    wl = _tfim32()
    assert_bitwise(*_device(wl), *_oracle("tfim32", wl), "table from the environment")


def test_clone_of_tuned_sum_bitwise():
    wl = _tfim32()
    pc = PackedCircuit(wl["circuit"], wl["thetas"], 32)
    s = HIPPauliSum(32)
    s.tune(merge_items=2, perm_items=4, grid_per_cu=1, lds_runs=False)
    k, c = wl["obs"].to_arrays(pc.W, 1)
    s.upload(k, c)
    c2 = s.clone()
    pc.run(c2, max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"])
    assert_bitwise(*c2.download(), *_oracle("tfim32", wl), "clone")


def test_installed_table_loads_for_every_class():
    import os
    path = os.path.join(os.path.dirname(_lib.__file__), "pphip_tune.txt")
    if not os.path.exists(path):
        pytest.skip("no tuning table installed beside the library")
    for nq, B in [(30, 1), (60, 1), (120, 1), (60, 32), (60, 64), (60, 2), (60, 100)]:
        HIPPauliSum(nq, batch=B).load_tuning(path)


@pytest.mark.parametrize("name", list(_WL))
@pytest.mark.parametrize("defer,items", [(0, 2), (0, 3), (1, 2), (1, 3), (1, 4)])
def test_deferred_merge_stores_bitwise(name, defer, items):
    # the one-column merge with each tile's stores one tile behind (k_merge1d)
    # and without, at every tile width
    print(f"{name} merge_defer={defer} merge_items={items}", flush=True)
    wl = _WL[name]()
    got = _device(wl, merge_defer=defer, merge_items=items)
    assert_bitwise(*got, *_oracle(name, wl), f"{name} merge_defer={defer} merge_items={items}")


@pytest.mark.parametrize("defer", [0, 1])
@pytest.mark.parametrize("mode", ["forced_rerun", "dynamic_tiles", "unchained"])
def test_deferred_merge_tile_modes_bitwise(defer, mode):
    # atomic tile claims (every gate rerun, or always) and merges that list
    # nothing for the next rotation, with stores one tile behind and without
    print(f"merge_defer={defer} {mode}", flush=True)
    wl = _kicked127(4)
    opts = {"forced_rerun": {"force_rerun": True}, "dynamic_tiles": {"dynamic_tiles": True},
            "unchained": {"chain_tails": False}}[mode]
    B = 1
    pc = PackedCircuit(wl["circuit"], wl["thetas"], wl["nqubits"], batch=B)
    s = HIPPauliSum(wl["nqubits"], batch=B, capacity=1024)
    s.tune(lds_runs=False, merge_defer=defer, merge_items=2, **opts)
    k, c = wl["obs"].to_arrays(pc.W, B)
    s.upload(k, c)
    pc.run(s, max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"])
    assert_bitwise(*s.download(), *_oracle("kicked127_4", wl), f"merge_defer={defer} {mode}")
    if mode == "forced_rerun":
        assert s.profile()["fallbacks"] > 0, "no gate was sent to the atomic-claim rerun"


@pytest.mark.parametrize("B", [5, 64])
@pytest.mark.parametrize("mode", ["forced_rerun", "dynamic_tiles", "tiny_spin_limit", "unchained"])
def test_batched_tile_modes_bitwise(B, mode):
    # the batched merge (stores one tile behind its merge paths) under atomic
    # tile claims, reruns and merges that list nothing for the next rotation
    print(f"B={B} {mode}", flush=True)
    wl = _heis(B, noise=True)
    opts = {"forced_rerun": {"force_rerun": True}, "dynamic_tiles": {"dynamic_tiles": True},
            "tiny_spin_limit": {"spin_limit": 1}, "unchained": {"chain_tails": False}}[mode]
    pc = PackedCircuit(wl["circuit"], wl["thetas"], wl["nqubits"], batch=B)
    s = HIPPauliSum(wl["nqubits"], batch=B, capacity=1024)
    s.tune(lds_runs=False, batch_rows=16, **opts)
    k, c = wl["obs"].to_arrays(pc.W, B)
    s.upload(k, c)
    pc.run(s, max_weight=wl["max_weight"], min_abs_coeff=wl["min_abs_coeff"])
    assert_bitwise(*s.download(), *_oracle(f"heis{B}True", wl), f"B={B} {mode}")
    if mode == "forced_rerun":
        assert s.profile()["fallbacks"] > 0, "no gate was sent to the atomic-claim rerun"


def _tfim66():
    from pauliprop_hip import workloads
    return workloads.tfim_tilted_6x6_36q()


@pytest.mark.parametrize("name", ["kicked127", "ansatz48", "tfim66"])
@pytest.mark.parametrize("below", [0, 1 << 31])
def test_self_split_merge_bitwise(name, below):
    # every one-column merge searching its own tile splits (whatever the sum's
    # size; one workgroup per CU, so large sums span several rounds of tiles
    # and the merge searches the splits of its later tiles too), or none; the
    # 6x6 TFIM grows to 1.04 M terms with 128-bit keys
    print(f"{name} split_below={below}", flush=True)
    wl = {**_WL, "tfim66": _tfim66}[name]()
    got = _device(wl, split_below=below, grid_per_cu=1)
    assert_bitwise(*got, *_oracle(name, wl), f"{name} split_below={below}")
