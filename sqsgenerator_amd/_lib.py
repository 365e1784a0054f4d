"""ctypes binding of libsqsgen_amd.so (declarations in include/sqsgen_amd.h).

The library is the product's only compute path: importing this module on a machine
without the built library raises, it never falls back to Python or CPU code.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libsqsgen_amd.so"

OK = 0
ERR_INVALID_ARGUMENT, ERR_OUT_OF_RANGE, ERR_DOMAIN, ERR_RUNTIME = 1, 2, 3, 4
ERR_DEVICE, ERR_NO_MEMORY, ERR_CANCELLED = 5, 6, 7

ITERATION_RANDOM, ITERATION_SYSTEMATIC = 0, 1
SUBLATTICE_INTERACT, SUBLATTICE_SPLIT = 0, 1
KERNEL_AUTO, KERNEL_SPARSE_WAVE, KERNEL_SPARSE_LANE, KERNEL_DENSE_MFMA, KERNEL_DENSE_PAIRED = \
    0, 1, 2, 3, 4

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
f32p = C.POINTER(C.c_float)


class SqsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


class DeviceError(SqsgError):
    pass


class CancelledError(SqsgError):
    pass


class Config(C.Structure):
    _fields_ = [
        ("num_sites", C.c_uint32), ("lattice", f64p), ("frac_coords", f64p), ("species", u8p),
        ("sublattice_mode", C.c_int32), ("iteration_mode", C.c_int32),
        ("num_sublattices", C.c_uint32), ("site_offsets", u32p), ("sites", u32p),
        ("comp_offsets", u32p), ("comp_species", u8p), ("comp_amounts", u32p),
        ("radii_offsets", u32p), ("radii", f64p), ("shell_offsets", u32p), ("shells", u32p),
        ("shell_weights", f64p), ("num_species", u32p), ("prefactors", f64p),
        ("pair_weights", f64p), ("target_objective", f64p), ("iterations", C.c_uint64),
        ("chunk_size", C.c_uint64), ("seeds", u64p), ("seed_present", u8p),
        ("keep", C.c_uint64), ("max_results_per_objective", C.c_int64),
        ("precision", C.c_int32),
    ]


class ProblemInfo(C.Structure):
    _fields_ = [("num_sites", C.c_uint32), ("num_species", C.c_uint32),
                ("num_shells", C.c_uint32), ("num_pairs", C.c_uint32),
                ("num_bounds", C.c_uint32), ("bounds", C.c_uint32 * 16),
                ("species_z", C.c_uint8 * 64), ("shell_ids", C.c_uint32 * 64)]


class Statistics(C.Structure):
    _fields_ = [("finished", C.c_uint64), ("working", C.c_uint64), ("best_rank", C.c_uint64),
                ("best_objective", C.c_double), ("timing_total_ns", C.c_uint64),
                ("timing_chunk_setup_ns", C.c_uint64), ("timing_loop_ns", C.c_uint64),
                ("timing_comm_ns", C.c_uint64)]


class BatchReport(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("candidates", C.c_uint64), ("launches", C.c_uint64),
                ("work_items", C.c_uint32)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_uint32),
                ("slice_candidates", C.c_uint64)]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.POINTER(Statistics), C.c_void_p)

# name -> (restype, argtypes); every entry point of include/sqsgen_amd.h
SIGNATURES = {
    "sqsg_last_error": (C.c_char_p, []),
    "sqsg_version": (C.c_int, []),
    "sqsg_device_info": (C.c_int, [C.c_char_p, C.c_size_t, C.POINTER(C.c_int)]),
    "sqsg_device_synchronize": (C.c_int, []),
    "sqsg_distance_matrix": (C.c_int, [f64p, f64p, C.c_uint32, f64p]),
    "sqsg_distances_naive": (C.c_int, [f64p, C.c_uint32, C.c_double, C.c_double, f64p,
                                       C.c_uint32, u32p]),
    "sqsg_distances_histogram": (C.c_int, [f64p, C.c_uint32, C.c_double, C.c_double, f64p,
                                           C.c_uint32, u32p]),
    "sqsg_shell_matrix": (C.c_int, [f64p, C.c_uint32, f64p, C.c_uint32, C.c_double,
                                    C.c_double, u32p]),
    "sqsg_prefactors": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, u8p, f64p]),
    "sqsg_supercell": (C.c_int, [f64p, f64p, u8p, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.c_uint32, f64p, f64p, u8p]),
    "sqsg_distance_matrix_f32": (C.c_int, [f32p, f32p, C.c_uint32, f32p]),
    "sqsg_distances_naive_f32": (C.c_int, [f32p, C.c_uint32, C.c_float, C.c_float, f32p,
                                           C.c_uint32, u32p]),
    "sqsg_distances_histogram_f32": (C.c_int, [f32p, C.c_uint32, C.c_float, C.c_float, f32p,
                                               C.c_uint32, u32p]),
    "sqsg_shell_matrix_f32": (C.c_int, [f32p, C.c_uint32, f32p, C.c_uint32, C.c_float,
                                        C.c_float, u32p]),
    "sqsg_prefactors_f32": (C.c_int, [u32p, C.c_uint32, u32p, C.c_uint32, u8p, f32p]),
    "sqsg_supercell_f32": (C.c_int, [f32p, f32p, u8p, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, f32p, f32p, u8p]),
    "sqsg_num_permutations": (C.c_int, [u8p, C.c_uint32, C.c_char_p, C.c_size_t]),
    "sqsg_rank_permutation": (C.c_int, [u8p, C.c_uint32, C.c_char_p, C.c_size_t]),
    "sqsg_unrank_permutation": (C.c_int, [u8p, C.c_uint32, C.c_char_p]),
    "sqsg_job_create": (C.c_int, [C.POINTER(Config), C.POINTER(C.c_void_p)]),
    "sqsg_job_destroy": (None, [C.c_void_p]),
    "sqsg_job_num_problems": (C.c_int, [C.c_void_p, u32p]),
    "sqsg_job_problem_info": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(ProblemInfo)]),
    "sqsg_job_problem_arrays": (C.c_int, [C.c_void_p, C.c_uint32, u32p, u8p, u32p, f64p, f64p,
                                          f64p]),
    "sqsg_score_batch": (C.c_int, [C.c_void_p, C.c_uint32, u8p, C.c_uint32, i32p, f64p, f64p,
                                   C.c_int]),
    "sqsg_score_batch_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                          C.c_void_p]),
    "sqsg_job_set_kernel": (C.c_int, [C.c_void_p, C.c_int]),
    "sqsg_job_set_launch_size": (C.c_int, [C.c_void_p, C.c_uint64]),
    "sqsg_run": (C.c_int, [C.c_void_p, PROGRESS_FN, C.c_void_p]),
    "sqsg_run_noalloc": (C.c_int, [C.c_void_p, PROGRESS_FN, C.c_void_p]),
    "sqsg_cancel": (None, [C.c_void_p]),
    "sqsg_job_statistics": (C.c_int, [C.c_void_p, C.POINTER(Statistics)]),
    "sqsg_job_timing": (C.c_int, [C.c_void_p, f64p, u64p, u64p]),
    "sqsg_result_count": (C.c_int, [C.c_void_p, u64p, u64p]),
    "sqsg_result_get": (C.c_int, [C.c_void_p, C.c_uint64, f64p, u8p, f64p, f64p, C.c_char_p,
                                  C.c_size_t, u64p]),
    "sqsg_result_species_packed": (C.c_int, [C.c_void_p, C.c_uint64, u8p]),
    "sqsg_result_levels": (C.c_int, [C.c_void_p, f64p, u64p]),
    "sqsg_result_bulk": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, f64p, u64p, u8p, u64p]),
    "sqsg_result_sro_bulk": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint64, f64p]),
    "sqsg_batch_run": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_int,
                                 C.POINTER(BatchReport)]),
    "sqsg_autotune": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), f64p]),
    "sqsg_job_kernel_shape": (C.c_int, [C.c_void_p, C.c_char_p, C.c_size_t]),
    "sqsg_job_set_pruning": (C.c_int, [C.c_void_p, C.c_int]),
    "sqsg_job_set_dense_launch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64]),
    "sqsg_job_dense_launch": (C.c_int, [C.c_void_p, u32p, u32p, u32p, u64p]),
    "sqsg_server_create": (C.c_int, [C.POINTER(ServerOptions), C.POINTER(C.c_void_p)]),
    "sqsg_server_submit": (C.c_int, [C.c_void_p, C.c_void_p, u64p]),
    "sqsg_server_step": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sqsg_server_poll": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_int), f64p, f64p,
                                   f64p]),
    "sqsg_server_destroy": (None, [C.c_void_p]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the library; raises when it has not been built (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m sqsgenerator_amd.build` "
                "(the HIP library is the only compute path)")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code == OK:
        return
    msg = lib().sqsg_last_error().decode()
    if code == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if code == ERR_OUT_OF_RANGE:
        raise IndexError(msg)
    if code == ERR_DOMAIN:
        raise ValueError(msg)
    if code == ERR_DEVICE:
        raise DeviceError(code, msg)
    if code == ERR_NO_MEMORY:
        raise MemoryError(msg)
    if code == ERR_CANCELLED:
        raise CancelledError(code, msg)
    raise SqsgError(code, msg)


def ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)
