"""pauliprop_hip -- MI355X backend for PauliPropagation.jl's propagate() hot path.

The user-facing names mirror the reference (Julia's `rxlayer!` is `rxlayer`
here); every propagation step runs in libpphip.so on the GPU.
"""
from .paulis import PauliString, PauliSum, symboltoint, key_words, countweight, containsXorY, getpauli
from .gates import (PauliRotation, ImaginaryPauliRotation, TGate, CliffordGate, DepolarizingNoise, PauliXNoise,
                    PauliYNoise, PauliZNoise, DephasingNoise, PauliXDamping, PauliYDamping, PauliZDamping,
                    AmplitudeDampingNoise, FrozenGate, freeze, countparameters, toheisenberg, toschrodinger,
                    transposecliffordmap, pack_circuit, pp_gate, CLIFFORD_MAP)
from .circuits import (bricklayertopology, staircasetopology, rectangletopology, heavyhextopology,
                       ibmeagletopology, rxlayer, rylayer, rzlayer, rxxlayer, ryylayer, rzzlayer,
                       depolarizinglayer, tfitrottercircuit, tiltedtfitrottercircuit, heisenbergtrottercircuit,
                       hardwareefficientcircuit, efficientsu2circuit)
from .hipsum import (HIPPauliSum, PackedCircuit, propagate, overlapwithzero, overlapwithplus,
                     overlapwithcomputational, overlapwithmaxmixed, overlapwithpaulisum, getcoeff, scalarproduct,
                     run_heisenberg, adjointgradient, rewindgradient, mcpropagate, resample, truncate,
                     maxabscoeff,
                     semideterministic_systematic_resample, systematic_resample, multinomial_resample)

__version__ = "0.1.0"
