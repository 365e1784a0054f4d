"""ctypes front-end of the C oracle (oracle/sqsgen_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker.  Imported by tests/, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of ``bench.py``;
never by the product package.  Each wrapper names the reference function it
restates (file:line in dgehringer/sqsgenerator v0.5.6).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
REF_PATH = HERE / "_ref" / "librapidhash_ref.so"

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
f64p = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    lib.or_rapid_mix.restype = C.c_uint64
    lib.or_rapid_mix.argtypes = [C.c_uint64, C.c_uint64]
    lib.or_rapidrand.restype = C.c_uint64
    lib.or_rapidrand.argtypes = [u64p]
    lib.or_random_bounded.restype = C.c_uint32
    lib.or_random_bounded.argtypes = [C.c_uint32, u64p]
    lib.or_shuffle_random.argtypes = [u8p, u32p, C.c_uint32, u64p]
    lib.or_num_permutations.argtypes = [u8p, C.c_uint32, u64p, u64p]
    lib.or_rank_permutation.argtypes = [u8p, C.c_uint32, u64p, u64p]
    lib.or_unrank_permutation.argtypes = [u8p, C.c_uint32, C.c_uint64, C.c_uint64]
    lib.or_next_permutation.argtypes = [u8p, C.c_uint32]
    lib.or_distance_matrix.argtypes = [f64p, f64p, C.c_uint32, f64p]
    lib.or_distances_naive.restype = C.c_uint32
    lib.or_distances_naive.argtypes = [f64p, C.c_uint32, C.c_double, C.c_double, f64p, C.c_uint32]
    lib.or_distances_histogram.argtypes = [f64p, C.c_uint32, C.c_double, C.c_double, f64p,
                                           C.c_uint32, u32p]
    lib.or_shell_matrix.argtypes = [f64p, C.c_uint32, f64p, C.c_uint32, C.c_double, C.c_double,
                                    u32p]
    lib.or_compute_prefactors.restype = C.c_uint32
    lib.or_compute_prefactors.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, u8p, f64p]
    lib.or_supercell.argtypes = [f64p, f64p, u8p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 f64p, f64p, u8p]
    lib.or_scaled_pair_weights.argtypes = [f64p, f64p, C.c_uint32, C.c_uint32]
    lib.or_pairs.restype = C.c_uint32
    lib.or_pairs.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, u64p, C.c_uint32]
    lib.or_count_bonds.argtypes = [u64p, C.c_uint32, u8p, C.c_uint32, C.c_uint32, u64p]
    lib.or_compute_objective.restype = C.c_double
    lib.or_compute_objective.argtypes = [u64p, f64p, f64p, f64p, C.c_uint32, C.c_uint32, f64p]
    f32p = C.POINTER(C.c_float)
    lib.or_distance_matrix_f32.argtypes = [f32p, f32p, C.c_uint32, f32p]
    lib.or_distance_matrix_f32.restype = None
    lib.or_shell_matrix_f32.argtypes = [f32p, C.c_uint32, f32p, C.c_uint32, C.c_float,
                                        C.c_float, u32p]
    lib.or_shell_matrix_f32.restype = None
    lib.or_compute_prefactors_f32.argtypes = [u32p, C.c_uint32, u32p, C.c_uint32, u8p, f32p]
    lib.or_compute_prefactors_f32.restype = C.c_uint32
    lib.or_scaled_pair_weights_f32.argtypes = [f32p, f32p, C.c_uint32, C.c_uint32]
    lib.or_scaled_pair_weights_f32.restype = None
    lib.or_compute_objective_f32.argtypes = [u64p, f32p, f32p, f32p, C.c_uint32, C.c_uint32,
                                             f32p]
    lib.or_compute_objective_f32.restype = C.c_float
    lib.or_run.argtypes = [C.POINTER(RunConfig), C.POINTER(RunStats)]
    lib.or_result_count.restype = C.c_uint64
    lib.or_result_get.argtypes = [C.c_uint64, f64p, u8p, f64p]
    return lib


class Sublattice(C.Structure):
    _fields_ = [("num_sites", C.c_uint32), ("num_species", C.c_uint32),
                ("num_shells", C.c_uint32), ("num_pairs", C.c_uint32),
                ("pairs", u64p), ("species_packed", u8p), ("bounds", u32p),
                ("num_bounds", C.c_uint32), ("prefactors", f64p), ("pair_weights", f64p),
                ("target", f64p), ("seed", C.c_uint64)]


class RunConfig(C.Structure):
    _fields_ = [("systematic", C.c_int), ("split", C.c_int), ("num_sublattices", C.c_uint32),
                ("sublattices", C.POINTER(Sublattice)), ("iterations", C.c_uint64),
                ("chunk_size", C.c_uint64), ("keep", C.c_uint64),
                ("max_results_per_objective", C.c_int64), ("num_threads", C.c_uint32),
                ("shared_rng", C.c_int), ("collect", C.c_int), ("counter_stream", C.c_int)]


class RunStats(C.Structure):
    _fields_ = [("finished", C.c_uint64), ("best_objective", C.c_double),
                ("num_results", C.c_uint64), ("num_objectives", C.c_uint64),
                ("seconds", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def ref_lib() -> C.CDLL | None:
    """oracle/_ref build of the reference's vendored rapidhash.h, when present."""
    if not REF_PATH.exists():
        return None
    r = C.CDLL(str(REF_PATH))
    r.ref_rapid_mix.restype = C.c_uint64
    r.ref_rapid_mix.argtypes = [C.c_uint64, C.c_uint64]
    r.ref_rapid_secret.restype = C.c_uint64
    r.ref_rapid_secret.argtypes = [C.c_int]
    return r


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


# -- RNG -------------------------------------------------------------------------------------

def rapid_mix(a: int, b: int) -> int:
    """rapidhash.h:209-212"""
    return lib().or_rapid_mix(a, b)


def rapidrand_stream(seed: int, count: int) -> np.ndarray:
    """shuffle.h:23-26, ``count`` successive draws"""
    st = C.c_uint64(seed)
    return np.array([lib().or_rapidrand(C.byref(st)) for _ in range(count)], dtype=np.uint64)


def random_bounded(rng: int, seed: int) -> tuple[int, int]:
    """shuffle.h:28-32; returns (value, new state)"""
    st = C.c_uint64(seed)
    v = lib().or_random_bounded(rng, C.byref(st))
    return v, st.value


def shuffle_random(conf: np.ndarray, bounds, seed: int) -> tuple[np.ndarray, int]:
    """shuffle.h:39-50 (Fisher-Yates per window); returns (shuffled copy, new state)"""
    c = np.ascontiguousarray(conf, dtype=np.uint8).copy()
    b = np.ascontiguousarray(np.asarray(bounds, dtype=np.uint32).reshape(-1))
    st = C.c_uint64(seed)
    lib().or_shuffle_random(_p(c, u8p), _p(b, u32p), len(b) // 2, C.byref(st))
    return c, st.value


# -- permutations ----------------------------------------------------------------------------

def num_permutations(conf) -> int:
    """permutation.h:17-32"""
    c = np.ascontiguousarray(conf, dtype=np.uint8)
    hi, lo = C.c_uint64(), C.c_uint64()
    if lib().or_num_permutations(_p(c, u8p), len(c), C.byref(hi), C.byref(lo)):
        raise OverflowError("multinomial exceeds 128 bits")
    return (hi.value << 64) | lo.value


def rank_permutation(conf) -> int:
    """permutation.h:34-50"""
    c = np.ascontiguousarray(conf, dtype=np.uint8)
    hi, lo = C.c_uint64(), C.c_uint64()
    lib().or_rank_permutation(_p(c, u8p), len(c), C.byref(hi), C.byref(lo))
    return (hi.value << 64) | lo.value


def unrank_permutation(conf, rank: int) -> np.ndarray:
    """permutation.h:52-78"""
    c = np.ascontiguousarray(conf, dtype=np.uint8).copy()
    rc = lib().or_unrank_permutation(_p(c, u8p), len(c), rank >> 64, rank & ((1 << 64) - 1))
    if rc == 2:
        raise IndexError("The rank is larger than the total number of permutations")
    if rc:
        raise OverflowError("multinomial exceeds 128 bits")
    return c


def next_permutation(conf) -> tuple[np.ndarray, bool]:
    """permutation.h:80-102"""
    c = np.ascontiguousarray(conf, dtype=np.uint8).copy()
    ok = lib().or_next_permutation(_p(c, u8p), len(c))
    return c, bool(ok)


# -- geometry --------------------------------------------------------------------------------

def distance_matrix(lattice, frac) -> np.ndarray:
    """structure.h:34-69"""
    L = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    f = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
    n = len(f)
    out = np.empty((n, n), dtype=np.float64)
    lib().or_distance_matrix(_p(L, f64p), _p(f, f64p), n, _p(out, f64p))
    return out


def distances_naive(dm, atol=np.finfo(float).eps, rtol=1e-9) -> np.ndarray:
    """structure.h:195-207"""
    d = np.ascontiguousarray(dm, dtype=np.float64)
    n = d.shape[0]
    out = np.empty(n * n + 1, dtype=np.float64)
    m = lib().or_distances_naive(_p(d, f64p), n, atol, rtol, _p(out, f64p), len(out))
    return out[:m].copy()


def distances_histogram(dm, bin_width=0.05, peak_isolation=0.25) -> np.ndarray:
    """structure.h:209-258"""
    d = np.ascontiguousarray(dm, dtype=np.float64)
    n = d.shape[0]
    out = np.empty(n * n + 2, dtype=np.float64)
    cnt = C.c_uint32()
    if lib().or_distances_histogram(_p(d, f64p), n, bin_width, peak_isolation, _p(out, f64p),
                                    len(out), C.byref(cnt)):
        raise ValueError("Not enough edges to create a histogram, please increase the bin width")
    return out[:cnt.value].copy()


def shell_matrix(dm, radii, atol=np.finfo(float).eps, rtol=1e-9) -> np.ndarray:
    """structure.h:71-100"""
    d = np.ascontiguousarray(dm, dtype=np.float64)
    r = np.ascontiguousarray(radii, dtype=np.float64)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    lib().or_shell_matrix(_p(d, f64p), n, _p(r, f64p), len(r), atol, rtol, _p(out, u32p))
    return out


def compute_prefactors(sm, shells, species_z) -> np.ndarray:
    """structure.h:136-192 (shells ascending)"""
    s = np.ascontiguousarray(sm, dtype=np.uint32)
    sh = np.ascontiguousarray(sorted(shells), dtype=np.uint32)
    z = np.ascontiguousarray(species_z, dtype=np.uint8)
    K = len(set(z.tolist()))
    out = np.zeros((len(sh), K, K), dtype=np.float64)
    lib().or_compute_prefactors(_p(s, u32p), s.shape[0], _p(sh, u32p), len(sh), _p(z, u8p),
                                _p(out, f64p))
    return out


def supercell(lattice, frac, species, a, b, c):
    """structure.h:335-366"""
    L = np.ascontiguousarray(lattice, dtype=np.float64).reshape(3, 3)
    f = np.ascontiguousarray(frac, dtype=np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(species, dtype=np.uint8)
    n = len(f)
    m = n * a * b * c
    Lo = np.empty((3, 3)); fo = np.empty((m, 3)); zo = np.empty(m, dtype=np.uint8)
    lib().or_supercell(_p(L, f64p), _p(f, f64p), _p(z, u8p), n, a, b, c, _p(Lo, f64p),
                       _p(fo, f64p), _p(zo, u8p))
    return Lo, fo, zo


def scaled_pair_weights(pw, shell_weights_sorted) -> np.ndarray:
    """optimization.h:17-33"""
    w = np.ascontiguousarray(pw, dtype=np.float64).copy()
    sw = np.ascontiguousarray(shell_weights_sorted, dtype=np.float64)
    S, K, _ = w.shape
    lib().or_scaled_pair_weights(_p(w, f64p), _p(sw, f64p), S, K)
    return w


def pairs(sm, shells) -> np.ndarray:
    """structure.h:434-451 -> (P, 3) uint64 records {i, j, packed shell}"""
    s = np.ascontiguousarray(sm, dtype=np.uint32)
    sh = np.ascontiguousarray(sorted(shells), dtype=np.uint32)
    n = s.shape[0]
    cap = n * (n - 1) // 2 + 1
    out = np.empty((cap, 3), dtype=np.uint64)
    P = lib().or_pairs(_p(s, u32p), n, _p(sh, u32p), len(sh), _p(out, u64p), cap)
    return out[:P].copy()


# -- hot path --------------------------------------------------------------------------------

def count_bonds(pair_list, species, S, K) -> np.ndarray:
    """optimization.h:74-77 -> (S, K, K) uint64, ordered (a = species[i<j])"""
    p = np.ascontiguousarray(pair_list, dtype=np.uint64)
    c = np.ascontiguousarray(species, dtype=np.uint8)
    out = np.zeros((S, K, K), dtype=np.uint64)
    lib().or_count_bonds(_p(p, u64p), len(p), _p(c, u8p), S, K, _p(out, u64p))
    return out


def compute_objective(bonds, prefactors, pair_weights, target):
    """optimization.h:79-103 -> (objective, sro cube)"""
    b = np.ascontiguousarray(bonds, dtype=np.uint64)
    S, K, _ = b.shape
    pf = np.ascontiguousarray(prefactors, dtype=np.float64)
    pw = np.ascontiguousarray(pair_weights, dtype=np.float64)
    tg = np.ascontiguousarray(target, dtype=np.float64)
    sro = np.zeros((S, K, K), dtype=np.float64)
    o = lib().or_compute_objective(_p(b, u64p), _p(pf, f64p), _p(pw, f64p), _p(tg, f64p), S, K,
                                   _p(sro, f64p))
    return o, sro


# -- driver ----------------------------------------------------------------------------------

class Problem:
    """Read-only inputs of one sublattice problem (optimization_config_data)."""

    def __init__(self, pairs_, species_packed, bounds, prefactors, pair_weights, target, seed=0):
        self.pairs = np.ascontiguousarray(pairs_, dtype=np.uint64).reshape(-1, 3)
        self.species = np.ascontiguousarray(species_packed, dtype=np.uint8)
        self.bounds = np.ascontiguousarray(np.asarray(bounds, dtype=np.uint32).reshape(-1))
        self.pref = np.ascontiguousarray(prefactors, dtype=np.float64)
        self.pw = np.ascontiguousarray(pair_weights, dtype=np.float64)
        self.tgt = np.ascontiguousarray(target, dtype=np.float64)
        self.seed = seed
        self.S, self.K, _ = self.pref.shape

    def struct(self) -> Sublattice:
        return Sublattice(len(self.species), self.K, self.S, len(self.pairs),
                          _p(self.pairs, u64p), _p(self.species, u8p), _p(self.bounds, u32p),
                          len(self.bounds) // 2, _p(self.pref, f64p), _p(self.pw, f64p),
                          _p(self.tgt, f64p), self.seed)


def run(problems, iterations, chunk_size=None, systematic=False, split=False, keep=1,
        max_results_per_objective=None, threads=1, shared_rng=True, collect=True,
        counter_stream=False):
    """sqs.h:299-467 worker loop; returns (stats, [(objective, species, sub_objectives)]).
    counter_stream: Monte Carlo candidate t is the sorted configuration shuffled with draws
    [t * D, (t + 1) * D) of each sublattice's seed stream (the device path's candidate
    definition) instead of the reference's cumulative shuffles"""
    arr = (Sublattice * len(problems))(*[p.struct() for p in problems])
    cfg = RunConfig(int(systematic), int(split), len(problems), arr, iterations,
                    chunk_size or min(iterations, 100000), keep,
                    -1 if max_results_per_objective is None else max_results_per_objective,
                    threads, int(shared_rng), int(collect), int(counter_stream))
    st = RunStats()
    if lib().or_run(C.byref(cfg), C.byref(st)):
        raise ValueError("invalid run configuration")
    results = []
    if collect:
        blob = sum(len(p.species) for p in problems)
        sp = np.empty(blob, dtype=np.uint8)
        so = np.empty(len(problems), dtype=np.float64)
        o = C.c_double()
        for r in range(lib().or_result_count()):
            lib().or_result_get(r, C.byref(o), _p(sp, u8p), _p(so, f64p))
            results.append((o.value, sp.copy(), so.copy()))
    return st, results


# ---- configuration<float>: the same functions with T = float ---------------------------

_f32p = C.POINTER(C.c_float)


def distance_matrix_f32(lattice, frac) -> np.ndarray:
    """structure.h:34-69, T = float"""
    L = np.ascontiguousarray(lattice, dtype=np.float32).reshape(3, 3)
    f = np.ascontiguousarray(frac, dtype=np.float32).reshape(-1, 3)
    n = len(f)
    out = np.empty((n, n), dtype=np.float32)
    lib().or_distance_matrix_f32(_p(L, _f32p), _p(f, _f32p), n, _p(out, _f32p))
    return out


def shell_matrix_f32(dm, radii, atol=float(np.finfo(np.float32).eps), rtol=1e-9) -> np.ndarray:
    """structure.h:71-100, T = float"""
    d = np.ascontiguousarray(dm, dtype=np.float32)
    r = np.ascontiguousarray(radii, dtype=np.float32)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    lib().or_shell_matrix_f32(_p(d, _f32p), n, _p(r, _f32p), len(r), atol, rtol, _p(out, u32p))
    return out


def compute_prefactors_f32(sm, shells, species_z) -> np.ndarray:
    """structure.h:136-192, T = float"""
    s = np.ascontiguousarray(sm, dtype=np.uint32)
    sh = np.ascontiguousarray(sorted(shells), dtype=np.uint32)
    z = np.ascontiguousarray(species_z, dtype=np.uint8)
    K = len(set(z.tolist()))
    out = np.zeros((len(sh), K, K), dtype=np.float32)
    lib().or_compute_prefactors_f32(_p(s, u32p), s.shape[0], _p(sh, u32p), len(sh), _p(z, u8p),
                                    _p(out, _f32p))
    return out


def scaled_pair_weights_f32(pw, shell_weights_sorted) -> np.ndarray:
    """optimization.h:17-33, T = float"""
    w = np.ascontiguousarray(pw, dtype=np.float32).copy()
    sw = np.ascontiguousarray(shell_weights_sorted, dtype=np.float32)
    S, K, _ = w.shape
    lib().or_scaled_pair_weights_f32(_p(w, _f32p), _p(sw, _f32p), S, K)
    return w


def compute_objective_f32(bonds, prefactors, pair_weights, target):
    """optimization.h:79-103, T = float -> (objective, sro cube) as float32"""
    b = np.ascontiguousarray(bonds, dtype=np.uint64)
    S, K, _ = b.shape
    pf = np.ascontiguousarray(prefactors, dtype=np.float32)
    pw = np.ascontiguousarray(pair_weights, dtype=np.float32)
    tg = np.ascontiguousarray(target, dtype=np.float32)
    sro = np.zeros((S, K, K), dtype=np.float32)
    o = lib().or_compute_objective_f32(_p(b, u64p), _p(pf, _f32p), _p(pw, _f32p), _p(tg, _f32p),
                                       S, K, _p(sro, _f32p))
    return np.float32(o), sro
