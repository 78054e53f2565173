"""ctypes binding of libpphip.so (include/pphip.h).

The library must be built in-tree (`__graft_entry__.build()` or
`make -C paulipropagation.jl_amd/csrc`); a missing library raises at load
time -- there is no host fallback.
"""
from __future__ import annotations

import ctypes
import os

from .gates import pp_gate

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libpphip.so")

PP_OK, PP_ERR_ARG, PP_ERR_NOMEM, PP_ERR_HIP, PP_ERR_UNSUPPORTED, PP_ERR_CAPACITY, PP_ERR_STATE = range(7)


class PPError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[pphip {code}] {msg}")
        self.code = code


class pp_gate_stats(ctypes.Structure):
    _fields_ = [("n_in", ctypes.c_int64), ("n_touched", ctypes.c_int64), ("n_out", ctypes.c_int64)]


KERNEL_KINDS = ("branch", "tail_permute", "merge", "noise", "overlap")


class pp_profile(ctypes.Structure):
    _fields_ = [("gates", ctypes.c_int64), ("term_applications", ctypes.c_int64), ("fallbacks", ctypes.c_int64),
                ("launches", ctypes.c_int64 * 5), ("milliseconds", ctypes.c_double * 5),
                ("algorithmic_bytes", ctypes.c_double * 5), ("identity_gates", ctypes.c_int64),
                ("persistent_gates", ctypes.c_int64), ("graph_launches", ctypes.c_int64),
                ("graph_captures", ctypes.c_int64)]


_lib = None

# (name, restype, argtypes) of every symbol the header declares
_P = ctypes.c_void_p
_U64P = ctypes.POINTER(ctypes.c_uint64)
_DP = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)
_SZ = ctypes.c_size_t
SIGNATURES = [
    ("pp_version", ctypes.c_int, []),
    ("pp_last_error", ctypes.c_char_p, []),
    ("pp_create", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _SZ, ctypes.POINTER(_P)]),
    ("pp_destroy", ctypes.c_int, [_P]),
    ("pp_clone", ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    ("pp_info", ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    ("pp_upload", ctypes.c_int, [_P, _U64P, _DP, _SZ]),
    ("pp_download", ctypes.c_int, [_P, _U64P, _DP, _SZ, ctypes.POINTER(_SZ)]),
    ("pp_load_merge", ctypes.c_int, [_P, _U64P, _DP, _SZ, ctypes.c_int, ctypes.c_double]),
    ("pp_apply_rotation", ctypes.c_int, [_P, _U64P, _DP, _DP, ctypes.c_int, ctypes.c_double,
                                         ctypes.POINTER(pp_gate_stats)]),
    ("pp_apply_pauli_noise", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_uint32, _DP, ctypes.c_int, ctypes.c_double,
                                            ctypes.POINTER(pp_gate_stats)]),
    ("pp_apply_imaginary_rotation", ctypes.c_int, [_P, _U64P, _DP, _DP, ctypes.c_int, ctypes.c_int,
                                                    ctypes.c_double, ctypes.POINTER(pp_gate_stats)]),
    ("pp_apply_amplitude_damping", ctypes.c_int, [_P, ctypes.c_int, _DP, ctypes.c_int, ctypes.c_double,
                                                   ctypes.POINTER(pp_gate_stats)]),
    ("pp_apply_clifford", ctypes.c_int, [_P, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _U64P, _DP, ctypes.c_int,
                                          ctypes.c_double, ctypes.POINTER(pp_gate_stats)]),
    ("pp_run_circuit", ctypes.c_int, [_P, ctypes.POINTER(pp_gate), _SZ, _DP, _SZ, _U64P, _DP, _SZ, ctypes.c_int,
                                      ctypes.c_double, _I64P]),
    ("pp_adjoint", ctypes.c_int, [_P, ctypes.POINTER(pp_gate), _SZ, _DP, _SZ, _U64P, _DP, _SZ, ctypes.c_int,
                                  ctypes.c_double, _SZ, ctypes.c_int, _U64P, _DP, _DP]),
    ("pp_rewind", ctypes.c_int, [_P, ctypes.POINTER(pp_gate), _SZ, _DP, _SZ, _U64P, _DP, _SZ, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, _U64P, _DP, _DP]),
    ("pp_overlap_zero", ctypes.c_int, [_P, _DP]),
    ("pp_overlap", ctypes.c_int, [_P, ctypes.c_int, _U64P, _DP]),
    ("pp_overlap_noalloc", ctypes.c_int, [_P, ctypes.c_int, _U64P, _DP]),
    ("pp_get_coeff", ctypes.c_int, [_P, _U64P, _DP]),
    ("pp_scalar_product", ctypes.c_int, [_P, _P, _DP]),
    ("pp_max_abs_coeff", ctypes.c_int, [_P, _DP]),
    ("pp_truncate", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, _DP]),
    ("pp_resample", ctypes.c_int, [_P, ctypes.c_int, _SZ, ctypes.c_int, ctypes.c_double, _DP, ctypes.c_int,
                                   ctypes.c_double, ctypes.c_double, _DP]),
    ("pp_synchronize", ctypes.c_int, [_P]),
    ("pp_profile_control", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int]),
    ("pp_profile_read", ctypes.c_int, [_P, ctypes.POINTER(pp_profile)]),
    ("pp_tune", ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int64]),
    ("pp_load_tuning", ctypes.c_int, [_P, ctypes.c_char_p]),
]

PP_TUNE_SPIN_LIMIT, PP_TUNE_DYNAMIC_TILES, PP_TUNE_CHAIN_TAILS, PP_TUNE_SKIP_IDENTITY = 1, 2, 3, 4
PP_TUNE_ADJ_DEFER_BYTES = 5
PP_TUNE_MERGE_ITEMS, PP_TUNE_PERM_ITEMS, PP_TUNE_BATCH_ROWS, PP_TUNE_RADIX_BITS = 6, 7, 8, 9
PP_TUNE_GRID_PER_CU, PP_TUNE_RUN_CHUNK = 10, 11
PP_TUNE_FORCE_RERUN, PP_TUNE_LDS_RUNS, PP_TUNE_GRAPHS = 12, 13, 14
PP_TUNE_MERGE_DEFER, PP_TUNE_SPLIT_BELOW = 15, 16
TUNE_NAMES = {"merge_items": PP_TUNE_MERGE_ITEMS, "merge_defer": PP_TUNE_MERGE_DEFER, "split_below": PP_TUNE_SPLIT_BELOW,
              "perm_items": PP_TUNE_PERM_ITEMS, "batch_rows": PP_TUNE_BATCH_ROWS,
              "radix_bits": PP_TUNE_RADIX_BITS, "grid_per_cu": PP_TUNE_GRID_PER_CU, "run_chunk": PP_TUNE_RUN_CHUNK}
PP_OVERLAP_ZERO, PP_OVERLAP_PLUS, PP_OVERLAP_COMPUTATIONAL, PP_OVERLAP_MAXMIXED = 0, 1, 2, 3
PP_RESAMPLE_SEMIDETERMINISTIC, PP_RESAMPLE_SYSTEMATIC, PP_RESAMPLE_MULTINOMIAL = 0, 1, 2


def lib():
    """Load libpphip.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"libpphip.so not found at {LIB_PATH}; build it with __graft_entry__.build() "
                "or `make -C paulipropagation.jl_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != PP_OK:
        msg = lib().pp_last_error().decode(errors="replace")
        if rc == PP_ERR_ARG:
            raise ValueError(f"[pphip] {msg}")
        raise PPError(rc, msg)
