"""Topologies and circuit builders with the reference's gate order
(src/Circuits/topologies.jl, src/Circuits/builders.jl). Qubits are 1-based.

The circuits are Schrodinger-ordered lists; propagate() reverses them.
"""
from __future__ import annotations

from .gates import CliffordGate, FrozenGate, PauliRotation, DepolarizingNoise


# ------------------------------------------------------------------ topologies

def bricklayertopology(nqubits, periodic: bool = False):
    """Two sublayers, odd-even then even-odd pairs; with `periodic` the pair
    (last, first) closes the layer whose parity it matches
    (src/Circuits/topologies.jl:24-50)."""
    q = list(range(1, nqubits + 1)) if isinstance(nqubits, int) else list(nqubits)
    n = len(q)
    topo = []
    if n == 1:
        return topo
    if n == 2:
        return [(q[0], q[1])]
    for ii in range(0, n - 1, 2):
        topo.append((q[ii], q[ii + 1]))
    if periodic and q[-1] % 2 == 1:
        topo.append((q[-1], q[0]))
    for ii in range(1, n - 1, 2):
        topo.append((q[ii], q[ii + 1]))
    if periodic and q[-1] % 2 == 0:
        topo.append((q[-1], q[0]))
    return topo


def staircasetopology(nqubits: int, periodic: bool = False):
    """(i, i+1) for i = 1..n-1, plus (n, 1) when periodic (src/Circuits/topologies.jl:57-63)."""
    topo = [(i, i + 1) for i in range(1, nqubits)]
    if periodic and nqubits > 2:
        topo.append((nqubits, 1))
    return topo


def rectangletopology(nx: int, ny: int, periodic: bool = False):
    """Grid couplings visited row by row, the vertical pair before the
    horizontal one, then the periodic wrap-around pairs with duplicates and
    self-pairs removed (src/Circuits/topologies.jl:71-111)."""
    topo = []
    for jj in range(1, ny + 1):
        for ii in range(1, nx + 1):
            if jj <= ny - 1:
                topo.append(((jj - 1) * nx + ii, jj * nx + ii))
            if ii + 1 <= nx:
                topo.append(((jj - 1) * nx + ii, (jj - 1) * nx + ii + 1))
    if periodic:
        nq = nx * ny
        for ii in range(1, nx + 1):
            topo.append((ii, nq - nx + ii))
        for ii in range(0, ny):
            topo.append((ii * nx + 1, ii * nx + nx))
        seen = set()
        uniq = []
        for p in topo:
            if p not in seen:
                seen.add(p)
                uniq.append(p)
        topo = [p for p in uniq if p[0] != p[1]]
    return topo


def heavyhextopology(nrows: int = 7, ncols: int = 15):
    """Heavy-hex coupling graph built row by row.

    Rows of `ncols` qubits are chained left to right; the first row lacks its
    last column and the last row its first. Between rows r and r+1 a bridge
    qubit couples the two rows at columns 0, 4, 8, ... when r is even and at
    2, 6, 10, ... when r is odd. Qubits are numbered row by row, each row's
    bridges after it; edges are listed as sorted pairs in ascending order.
    nrows=7, ncols=15 gives the 127-qubit Eagle graph of
    src/Circuits/topologies.jl:185-200 (degree <= 3, 144 edges).
    """
    index = {}
    nxt = 1
    edges = []
    rows = []
    for r in range(nrows):
        cols = list(range(ncols))
        if r == 0:
            cols = cols[:-1]
        if r == nrows - 1:
            cols = cols[1:]
        row = {}
        for c in cols:
            row[c] = nxt
            index[(r, c)] = nxt
            nxt += 1
        rows.append(row)
        for a, b in zip(cols, cols[1:]):
            edges.append((row[a], row[b]))
        if r > 0:
            for c, bq in bridges_prev:
                edges.append((rows[r - 1][c], bq))
                edges.append((bq, row[c]))
        if r < nrows - 1:
            start = 0 if r % 2 == 0 else 2
            bridges_prev = []
            for c in range(start, ncols, 4):
                bridges_prev.append((c, nxt))
                nxt += 1
    edges = sorted(tuple(sorted(e)) for e in edges)
    return edges


# The 127-qubit IBM Eagle coupling list of the reference, kept edge for edge
# as data (src/Circuits/topologies.jl:185-200); heavyhextopology() rebuilds it.
ibmeagletopology = [
    (1, 2), (1, 15), (2, 3), (3, 4), (4, 5), (5, 6), (5, 16), (6, 7), (7, 8), (8, 9), (9, 10), (9, 17),
    (10, 11), (11, 12), (12, 13), (13, 14), (13, 18), (15, 19), (16, 23), (17, 27), (18, 31), (19, 20),
    (20, 21), (21, 22), (21, 34), (22, 23), (23, 24), (24, 25), (25, 26), (25, 35), (26, 27), (27, 28),
    (28, 29), (29, 30), (29, 36), (30, 31), (31, 32), (32, 33), (33, 37), (34, 40), (35, 44), (36, 48),
    (37, 52), (38, 39), (38, 53), (39, 40), (40, 41), (41, 42), (42, 43), (42, 54), (43, 44), (44, 45),
    (45, 46), (46, 47), (46, 55), (47, 48), (48, 49), (49, 50), (50, 51), (50, 56), (51, 52), (53, 57),
    (54, 61), (55, 65), (56, 69), (57, 58), (58, 59), (59, 60), (59, 72), (60, 61), (61, 62), (62, 63),
    (63, 64), (63, 73), (64, 65), (65, 66), (66, 67), (67, 68), (67, 74), (68, 69), (69, 70), (70, 71),
    (71, 75), (72, 78), (73, 82), (74, 86), (75, 90), (76, 77), (76, 91), (77, 78), (78, 79), (79, 80),
    (80, 81), (80, 92), (81, 82), (82, 83), (83, 84), (84, 85), (84, 93), (85, 86), (86, 87), (87, 88),
    (88, 89), (88, 94), (89, 90), (91, 95), (92, 99), (93, 103), (94, 107), (95, 96), (96, 97), (97, 98),
    (97, 110), (98, 99), (99, 100), (100, 101), (101, 102), (101, 111), (102, 103), (103, 104), (104, 105),
    (105, 106), (105, 112), (106, 107), (107, 108), (108, 109), (109, 113), (110, 115), (111, 119), (112, 123),
    (113, 127), (114, 115), (115, 116), (116, 117), (117, 118), (118, 119), (119, 120), (120, 121), (121, 122),
    (122, 123), (123, 124), (124, 125), (125, 126), (126, 127),
]


# ------------------------------------------------------------------ layers

def rxlayer(circuit, nqubits):
    circuit.extend(PauliRotation("X", i) for i in range(1, nqubits + 1))
    return circuit


def rylayer(circuit, nqubits):
    circuit.extend(PauliRotation("Y", i) for i in range(1, nqubits + 1))
    return circuit


def rzlayer(circuit, nqubits):
    circuit.extend(PauliRotation("Z", i) for i in range(1, nqubits + 1))
    return circuit


def rxxlayer(circuit, topology):
    circuit.extend(PauliRotation(["X", "X"], list(p)) for p in topology)
    return circuit


def ryylayer(circuit, topology):
    circuit.extend(PauliRotation(["Y", "Y"], list(p)) for p in topology)
    return circuit


def rzzlayer(circuit, topology):
    circuit.extend(PauliRotation(["Z", "Z"], list(p)) for p in topology)
    return circuit


def depolarizinglayer(circuit, nqubits, p):
    circuit.extend(DepolarizingNoise(i, p) for i in range(1, nqubits + 1))
    return circuit


# ------------------------------------------------------------------ circuits

def tfitrottercircuit(nqubits: int, nlayers: int, topology=None, start_with_ZZ: bool = True):
    """src/Circuits/builders.jl:20-54"""
    topology = bricklayertopology(nqubits) if topology is None else topology
    circ = []
    if nlayers == 0:
        return circ
    if start_with_ZZ:
        rzzlayer(circ, topology)
    for _ in range(nlayers - 1):
        rxlayer(circ, nqubits)
        rzzlayer(circ, topology)
    rxlayer(circ, nqubits)
    if not start_with_ZZ:
        rzzlayer(circ, topology)
    return circ


def tiltedtfitrottercircuit(nqubits: int, nlayers: int, topology=None):
    """src/Circuits/builders.jl:66-78"""
    topology = bricklayertopology(nqubits) if topology is None else topology
    circ = []
    for _ in range(nlayers):
        rzzlayer(circ, topology)
        rzlayer(circ, nqubits)
        rxlayer(circ, nqubits)
    return circ


def heisenbergtrottercircuit(nqubits: int, nlayers: int, topology=None):
    """src/Circuits/builders.jl:82-92"""
    topology = bricklayertopology(nqubits) if topology is None else topology
    circ = []
    for _ in range(nlayers):
        rxxlayer(circ, topology)
        ryylayer(circ, topology)
        rzzlayer(circ, topology)
    return circ


def hardwareefficientcircuit(nqubits: int, nlayers: int, topology=None):
    """src/Circuits/builders.jl:102-123"""
    topology = bricklayertopology(nqubits) if topology is None else topology
    circ = []
    for _ in range(nlayers):
        for ii in range(1, nqubits + 1):
            circ.append(PauliRotation("X", ii))
            circ.append(PauliRotation("Z", ii))
            circ.append(PauliRotation("X", ii))
        for pair in topology:
            circ.append(PauliRotation(["Y", "Y"], list(pair)))
    return circ


def efficientsu2circuit(nqubits: int, nlayers: int, topology=None):
    """src/Circuits/builders.jl:132-153"""
    topology = bricklayertopology(nqubits) if topology is None else topology
    circ = []
    for _ in range(nlayers):
        for ii in range(1, nqubits + 1):
            circ.append(PauliRotation("Y", ii))
            circ.append(PauliRotation("Z", ii))
        for pair in topology:
            circ.append(CliffordGate("CNOT", list(pair)))
    return circ


__all__ = [
    "bricklayertopology", "staircasetopology", "rectangletopology", "heavyhextopology", "ibmeagletopology",
    "rxlayer", "rylayer", "rzlayer", "rxxlayer", "ryylayer", "rzzlayer", "depolarizinglayer",
    "tfitrottercircuit", "tiltedtfitrottercircuit", "heisenbergtrottercircuit", "hardwareefficientcircuit",
    "efficientsu2circuit", "FrozenGate",
]
