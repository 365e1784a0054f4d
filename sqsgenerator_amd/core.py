"""User-facing API mirroring sqsgenerator's Python package.

``parse_config`` restates io/config/combined.h:192-295 (the subset on the optimisation
path: structure, composition, shell radii/weights, prefactors, pair weights, targets,
iterations, seeds, chunk size, keep, max_results_per_objective), ``optimize`` restates
python/sqsgenerator/_optimize.py:147-168 on top of the C ABI (include/sqsgen_amd.h), and
``SqsResultPack`` / ``SqsResult`` mirror the bindings of python/extension/sqsgen.cpp:252-335.
Parse errors are returned as ``ParseError`` values like the reference's parse_result.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass, field
from typing import Any, Callable, Iterator, Optional, Union

import numpy as np

from . import _lib as L
from .elements import Z_OF, symbol, to_z

KNOWN_KEYS = ("iteration_mode", "sublattice_mode", "seed", "structure", "composition",
              "shell_radii", "shell_weights", "prefactors", "pair_weights", "target_objective",
              "iterations", "chunk_size", "thread_config", "keep", "max_results_per_objective",
              "atol", "rtol", "prec", "bin_width", "peak_isolation")  # combined.h:17-36

ITERATIONS_DEFAULT = 500000    # combined.h:38
CHUNK_SIZE_DEFAULT = 100000    # combined.h:39
ATOL_DEFAULT, RTOL_DEFAULT = 1.0e-3, 1.0e-5            # shell_radii.h:19-20
BIN_WIDTH_DEFAULT, PEAK_ISOLATION_DEFAULT = 0.05, 0.25  # shell_radii.h:21-22
EPS = float(np.finfo(np.float64).eps)
EPS_F32 = float(np.finfo(np.float32).eps)

random, systematic = "random", "systematic"
interact, split = "interact", "split"
double, single = "double", "single"


class ParseErrorCode:  # io/parsing.h:19-26
    CODE_UNKNOWN = "unknown"
    CODE_NOT_FOUND = "not_found"
    CODE_TYPE_ERROR = "type_error"
    CODE_BAD_VALUE = "bad_value"
    CODE_BAD_ARGUMENT = "bad_argument"
    CODE_OUT_OF_RANGE = "out_of_range"


@dataclass
class ParseError:
    key: str
    msg: str
    code: str
    parameter: Optional[str] = None

    def __bool__(self) -> bool:
        return False


class _Fail(Exception):
    def __init__(self, key, code, msg):
        super().__init__(msg)
        self.err = ParseError(key, msg, code)


# ---- geometry wrappers (core/structure.h) ------------------------------------------------

def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


def _as_prec(x, single: bool):
    """a value of the configuration's type T (float values are kept as float64)"""
    return float(np.float32(x)) if single else float(x)


def distance_matrix(lattice, frac_coords, single: bool = False) -> np.ndarray:
    if single:  # structure<float>
        lat, frac = _f32(lattice).reshape(3, 3), _f32(frac_coords).reshape(-1, 3)
        n = len(frac)
        out = np.empty((n, n), dtype=np.float32)
        L.check(L.lib().sqsg_distance_matrix_f32(L.ptr(lat, L.f32p), L.ptr(frac, L.f32p), n,
                                                 L.ptr(out, L.f32p)))
        return out.astype(np.float64)
    lat, frac = _f64(lattice).reshape(3, 3), _f64(frac_coords).reshape(-1, 3)
    n = len(frac)
    out = np.empty((n, n))
    L.check(L.lib().sqsg_distance_matrix(L.ptr(lat, L.f64p), L.ptr(frac, L.f64p), n,
                                         L.ptr(out, L.f64p)))
    return out


def distances_naive(dm, atol=EPS, rtol=1e-9, single: bool = False) -> np.ndarray:
    if single:
        d = _f32(dm)
        n = d.shape[0]
        out = np.empty(n * n + 1, dtype=np.float32)
        cnt = C.c_uint32()
        L.check(L.lib().sqsg_distances_naive_f32(L.ptr(d, L.f32p), n, atol, rtol,
                                                 L.ptr(out, L.f32p), len(out), C.byref(cnt)))
        return out[:cnt.value].astype(np.float64)
    d = _f64(dm)
    n = d.shape[0]
    out = np.empty(n * n + 1)
    cnt = C.c_uint32()
    L.check(L.lib().sqsg_distances_naive(L.ptr(d, L.f64p), n, atol, rtol, L.ptr(out, L.f64p),
                                         len(out), C.byref(cnt)))
    return out[:cnt.value].copy()


def distances_histogram(dm, bin_width=BIN_WIDTH_DEFAULT,
                        peak_isolation=PEAK_ISOLATION_DEFAULT, single: bool = False) -> np.ndarray:
    if single:
        d = _f32(dm)
        n = d.shape[0]
        out = np.empty(n * n + 2, dtype=np.float32)
        cnt = C.c_uint32()
        L.check(L.lib().sqsg_distances_histogram_f32(L.ptr(d, L.f32p), n, bin_width,
                                                     peak_isolation, L.ptr(out, L.f32p),
                                                     len(out), C.byref(cnt)))
        return out[:cnt.value].astype(np.float64)
    d = _f64(dm)
    n = d.shape[0]
    out = np.empty(n * n + 2)
    cnt = C.c_uint32()
    L.check(L.lib().sqsg_distances_histogram(L.ptr(d, L.f64p), n, bin_width, peak_isolation,
                                             L.ptr(out, L.f64p), len(out), C.byref(cnt)))
    return out[:cnt.value].copy()


def shell_matrix(dm, radii, atol=EPS, rtol=1e-9, single: bool = False) -> np.ndarray:
    if single:
        d, r = _f32(dm), _f32(radii)
        n = d.shape[0]
        out = np.empty((n, n), dtype=np.uint32)
        L.check(L.lib().sqsg_shell_matrix_f32(L.ptr(d, L.f32p), n, L.ptr(r, L.f32p), len(r),
                                              atol, rtol, L.ptr(out, L.u32p)))
        return out
    d, r = _f64(dm), _f64(radii)
    n = d.shape[0]
    out = np.empty((n, n), dtype=np.uint32)
    L.check(L.lib().sqsg_shell_matrix(L.ptr(d, L.f64p), n, L.ptr(r, L.f64p), len(r), atol, rtol,
                                      L.ptr(out, L.u32p)))
    return out


def prefactors(sm, shells, species, single: bool = False) -> np.ndarray:
    s = np.ascontiguousarray(sm, dtype=np.uint32)
    sh = np.ascontiguousarray(sorted(shells), dtype=np.uint32)
    z = np.ascontiguousarray(species, dtype=np.uint8)
    K = len(set(z.tolist()))
    if single:
        outf = np.zeros((len(sh), K, K), dtype=np.float32)
        L.check(L.lib().sqsg_prefactors_f32(L.ptr(s, L.u32p), s.shape[0], L.ptr(sh, L.u32p),
                                            len(sh), L.ptr(z, L.u8p), L.ptr(outf, L.f32p)))
        return outf.astype(np.float64)
    out = np.zeros((len(sh), K, K))
    L.check(L.lib().sqsg_prefactors(L.ptr(s, L.u32p), s.shape[0], L.ptr(sh, L.u32p), len(sh),
                                    L.ptr(z, L.u8p), L.ptr(out, L.f64p)))
    return out


def num_permutations(conf) -> int:
    c = np.ascontiguousarray(conf, dtype=np.uint8)
    buf = C.create_string_buffer(4096)
    L.check(L.lib().sqsg_num_permutations(L.ptr(c, L.u8p), len(c), buf, 4096))
    return int(buf.value.decode())


def rank_permutation(conf) -> int:
    c = np.ascontiguousarray(conf, dtype=np.uint8)
    buf = C.create_string_buffer(4096)
    L.check(L.lib().sqsg_rank_permutation(L.ptr(c, L.u8p), len(c), buf, 4096))
    return int(buf.value.decode())


def unrank_permutation(conf, rank: int) -> np.ndarray:
    c = np.ascontiguousarray(conf, dtype=np.uint8).copy()
    L.check(L.lib().sqsg_unrank_permutation(L.ptr(c, L.u8p), len(c), str(int(rank)).encode()))
    return c


class Structure:
    """Lattice (rows a, b, c), fractional coordinates and atomic numbers."""

    def __init__(self, lattice, frac_coords, species, pbc=(True, True, True),
                 single: bool = False):
        # structure<float> (single) holds float values, kept as float64
        self.single = bool(single)
        cast = (lambda a: _f64(_f32(a))) if self.single else _f64
        self.lattice = cast(lattice).reshape(3, 3)
        self.frac_coords = cast(frac_coords).reshape(-1, 3)
        self.species = [to_z(s) for s in species]
        self.pbc = tuple(pbc)
        if len(self.species) != len(self.frac_coords) or not self.species:
            raise ValueError(
                "frac coords must have the same size as the species input and must not be empty")

    def __len__(self) -> int:
        return len(self.species)

    def __iter__(self) -> Iterator[tuple[int, np.ndarray, int]]:
        for i, (c, s) in enumerate(zip(self.frac_coords, self.species)):
            yield i, c, s

    @property
    def symbols(self) -> list[str]:
        return [symbol(z) for z in self.species]

    def supercell(self, a: int, b: int, c: int) -> "Structure":
        n = len(self)
        m = n * a * b * c
        z = np.ascontiguousarray(self.species, dtype=np.uint8)
        if self.single:
            lat, frac = _f32(self.lattice), _f32(self.frac_coords)
            lo, fo = np.empty((3, 3), np.float32), np.empty((m, 3), np.float32)
            zo = np.empty(m, dtype=np.uint8)
            L.check(L.lib().sqsg_supercell_f32(L.ptr(lat, L.f32p), L.ptr(frac, L.f32p),
                                               L.ptr(z, L.u8p), n, a, b, c, L.ptr(lo, L.f32p),
                                               L.ptr(fo, L.f32p), L.ptr(zo, L.u8p)))
            return Structure(lo, fo, zo.tolist(), self.pbc, True)
        lo, fo, zo = np.empty((3, 3)), np.empty((m, 3)), np.empty(m, dtype=np.uint8)
        L.check(L.lib().sqsg_supercell(L.ptr(self.lattice, L.f64p), L.ptr(self.frac_coords, L.f64p),
                                       L.ptr(z, L.u8p), n, a, b, c, L.ptr(lo, L.f64p),
                                       L.ptr(fo, L.f64p), L.ptr(zo, L.u8p)))
        return Structure(lo, fo, zo.tolist(), self.pbc)

    def distance_matrix(self) -> np.ndarray:
        return distance_matrix(self.lattice, self.frac_coords, self.single)

    def shell_matrix(self, radii, atol=None, rtol=1e-9) -> np.ndarray:
        if atol is None:  # numeric_limits<T>::epsilon()
            atol = EPS_F32 if self.single else EPS
        return shell_matrix(self.distance_matrix(), radii, atol, rtol, self.single)

    def sliced(self, indices) -> "Structure":
        idx = list(indices)
        return Structure(self.lattice, self.frac_coords[idx], [self.species[i] for i in idx],
                         self.pbc, self.single)

    def with_species(self, species) -> "Structure":
        return Structure(self.lattice, self.frac_coords, species, self.pbc, self.single)

    def bytes(self) -> bytes:
        """msgpack wire form (sqsgen.cpp:182, io/json.h:139-157)"""
        from . import pack_io as W
        return W.dumps(W.structure_doc(self))

    @staticmethod
    def from_bytes(data: bytes) -> "Structure":
        from . import pack_io as W
        return W.structure_from_doc(W.loads(bytes(data)))


# ---- configuration ------------------------------------------------------------------------

@dataclass
class Sublattice:
    sites: list[int]
    composition: dict[int, int]   # Z -> amount, ascending Z


@dataclass
class SqsConfiguration:
    """Mirror of core/config.h:32-48 configuration<T>, T = double or float (`prec`)."""
    sublattice_mode: str
    iteration_mode: str
    seed: list[Optional[int]]
    structure_config: dict
    composition: list[Sublattice]
    shell_radii: list[list[float]]
    shell_weights: list[dict[int, float]]
    prefactors: list[np.ndarray]
    pair_weights: list[np.ndarray]
    target_objective: list[np.ndarray]
    iterations: int
    chunk_size: int
    thread_config: list[int]
    keep: int
    max_results_per_objective: Optional[int]
    prec: str = double
    _structure: Structure = field(default=None, repr=False)

    def structure(self) -> Structure:
        return self._structure

    def bytes(self) -> bytes:
        """msgpack wire form (sqsgen.cpp:222, io/json.h:177-195)"""
        from . import pack_io as W
        return W.dumps(W.config_doc(self))

    @staticmethod
    def from_bytes(data: bytes) -> "SqsConfiguration":
        from . import pack_io as W
        return W.config_from_doc(W.loads(bytes(data)))

    def json(self) -> str:
        d = {
            "sublattice_mode": self.sublattice_mode, "iteration_mode": self.iteration_mode,
            "structure": self.structure_config,
            "composition": [dict({"sites": sl.sites}, **{symbol(z): a for z, a in
                                                          sl.composition.items()})
                            for sl in self.composition],
            "shell_radii": self.shell_radii,
            "shell_weights": [{str(k): v for k, v in w.items()} for w in self.shell_weights],
            "prefactors": [p.tolist() for p in self.prefactors],
            "pair_weights": [p.tolist() for p in self.pair_weights],
            "target_objective": [p.tolist() for p in self.target_objective],
            "iterations": self.iterations, "chunk_size": self.chunk_size,
            "thread_config": self.thread_config, "keep": self.keep,
            "max_results_per_objective": self.max_results_per_objective, "prec": self.prec,
        }
        if any(s is not None for s in self.seed):
            d["seed"] = self.seed
        return json.dumps(d)


def _mode(doc, key, allowed, default):
    v = doc.get(key, default)
    if hasattr(v, "name"):
        v = v.name
    v = str(v).lower()
    if v not in allowed:
        raise _Fail(key, ParseErrorCode.CODE_BAD_VALUE, f"Invalid {key} {v!r}")
    return v


def _parse_structure(doc, single: bool = False) -> tuple[Structure, dict]:
    sc = doc.get("structure")
    if sc is None:
        raise _Fail("structure", ParseErrorCode.CODE_NOT_FOUND, "You need to specify a structure")
    try:
        lat = _f64(sc["lattice"]).reshape(3, 3)
        coords = _f64(sc["coords"]).reshape(-1, 3)
        species = [to_z(s) for s in sc["species"]]
        cell = [int(x) for x in sc.get("supercell", [1, 1, 1])]
    except (KeyError, ValueError, TypeError) as e:
        raise _Fail("structure", ParseErrorCode.CODE_BAD_VALUE, f"invalid structure: {e}")
    if len(cell) != 3 or min(cell) < 1:
        raise _Fail("supercell", ParseErrorCode.CODE_BAD_VALUE,
                    "There must be at least one copy in each direction")
    try:
        base = Structure(lat, coords, species, single=single)
    except ValueError as e:
        raise _Fail("structure", ParseErrorCode.CODE_BAD_VALUE, str(e))
    st = base.supercell(*cell)
    return st, {"lattice": lat.tolist(), "coords": coords.tolist(), "species": species,
                "supercell": cell}


def _parse_sites(spec, species: list[int], remaining: list[int]) -> list[int]:
    """composition.h:26-86: index list, symbol or symbol list; default = remaining sites"""
    if spec is None:
        return list(remaining)
    if isinstance(spec, str):
        spec = [spec]
    spec = list(spec)
    if spec and all(isinstance(s, str) for s in spec):
        present = set(species)
        zs = set()
        for s in spec:
            if s not in Z_OF:
                raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE,
                            f'An atomic element with name "{s}" is not known to me')
            if Z_OF[s] not in present:
                raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE,
                            f'The specified structure does not contain a site with species "{s}"')
            zs.add(Z_OF[s])
        return [i for i, z in enumerate(species) if z in zs]
    idx = sorted({int(i) for i in spec})
    for i in idx:
        if not 0 <= i < len(species):
            raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE,
                        f"Invalid index {i}, the specified structure contains {len(species)} sites")
    if not idx:
        raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE, "You cannot specify an empty slice")
    return idx


def _parse_composition(doc, st: Structure, mode: str) -> list[Sublattice]:
    if "composition" not in doc:
        raise _Fail("composition", ParseErrorCode.CODE_NOT_FOUND,
                    "You need to specify a composition")
    comp = doc["composition"]
    entries = comp if isinstance(comp, (list, tuple)) else [comp]
    if not isinstance(comp, (list, tuple)) and mode == split:
        raise _Fail("composition", ParseErrorCode.CODE_BAD_VALUE,
                    'You have specified "split" mode but only specified one sublattice')
    subs: list[Sublattice] = []
    occupied: set[int] = set()
    for e in entries:
        if not isinstance(e, dict):
            raise _Fail("composition", ParseErrorCode.CODE_BAD_VALUE,
                        "the JSON value is not an object")
        remaining = [i for i in range(len(st)) if i not in occupied]
        if not remaining:
            raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE,
                        "There are no remaining species left for the sublattice")
        sites = _parse_sites(e.get("sites"), st.species, remaining)
        for s in sites:
            if s in occupied:
                raise _Fail("sites", ParseErrorCode.CODE_BAD_VALUE,
                            f"The site with index {s} is contained in more than one sublattice")
        distribution: dict[int, int] = {}
        for k, v in e.items():
            if k == "sites" or k not in Z_OF:
                continue
            amount = int(v)
            if amount < 0 or amount > len(sites):
                raise _Fail("composition", ParseErrorCode.CODE_BAD_VALUE,
                            f'You want to distribute {amount} "{k}" atoms on a sublattice with '
                            f"{len(sites)} sites")
            distribution[Z_OF[k]] = amount
        if not distribution:
            raise _Fail("composition", ParseErrorCode.CODE_OUT_OF_RANGE,
                        "You have not distributed any valid species on the (sub)lattice")
        if sum(distribution.values()) != len(sites):
            raise _Fail("sites", ParseErrorCode.CODE_OUT_OF_RANGE,
                        f"The total number of distributed atoms is {sum(distribution.values())} "
                        f"but the sublattice contains {len(sites)} sites")
        occupied.update(sites)
        subs.append(Sublattice(sites, dict(sorted(distribution.items()))))
    if not subs:
        raise _Fail("composition", ParseErrorCode.CODE_OUT_OF_RANGE,
                    "Could not parse a valid sublattice")
    return subs


def _apply_composition(st: Structure, subs: list[Sublattice]) -> Structure:
    """structure.h:390-399"""
    species = list(st.species)
    for sl in subs:
        it = iter(sl.sites)
        for z, amount in sl.composition.items():
            for _ in range(amount):
                species[next(it)] = z
    return st.with_species(species)


def _radii(doc, value, st: Structure) -> list[float]:
    """shell_radii.h:27-72"""
    if isinstance(value, str):
        mode = value.lower()
        dm = st.distance_matrix()
        t = st.single
        if mode == "naive":
            return distances_naive(dm, _as_prec(doc.get("atol", ATOL_DEFAULT), t),
                                   _as_prec(doc.get("rtol", RTOL_DEFAULT), t), t).tolist()
        if mode == "peak":
            try:
                return distances_histogram(
                    dm, _as_prec(doc.get("bin_width", BIN_WIDTH_DEFAULT), t),
                    _as_prec(doc.get("peak_isolation", PEAK_ISOLATION_DEFAULT), t), t).tolist()
            except ValueError as e:
                raise _Fail("shell_radii", ParseErrorCode.CODE_BAD_VALUE, str(e))
        raise _Fail("shell_radii", ParseErrorCode.CODE_BAD_VALUE,
                    'Invalid shell radii detection mode must be either "naive" or "peak"')
    radii = [_as_prec(r, st.single) for r in value]
    if not radii:
        raise _Fail("shell_radii", ParseErrorCode.CODE_OUT_OF_RANGE,
                    "You need to define at least one coordination shells")
    if any(r < 0 for r in radii):
        raise _Fail("shell_radii", ParseErrorCode.CODE_BAD_VALUE,
                    "You cannot specify a shell radius that is less than 0")
    if radii[0] != 0.0 and not abs(radii[0]) <= _as_prec(1e-5, st.single):
        radii.insert(0, 0.0)
    return radii


def _weights(value, num_radii: int, single: bool = False) -> dict[int, float]:
    """shell_weights.h:186-223 (defaults T(1) / T(i), shell_weights.h:75-84)"""
    if value is None:
        return {i: _as_prec(1.0 / i, single) for i in range(1, num_radii)}
    if not isinstance(value, dict):
        raise _Fail("shell_weights", ParseErrorCode.CODE_BAD_VALUE,
                    '"shell_weights" must be an object')
    w: dict[int, float] = {}
    for k, v in value.items():
        s = int(k)
        if s == 0:
            raise _Fail("shell_weights", ParseErrorCode.CODE_BAD_VALUE,
                        "There is no point in including self-interactions")
        if s >= num_radii:
            raise _Fail("shell_weights", ParseErrorCode.CODE_OUT_OF_RANGE,
                        f"There are {num_radii} shells but you specified a shell index of {s}")
        w[s] = _as_prec(v, single)
    if not w:
        raise _Fail("shell_weights", ParseErrorCode.CODE_OUT_OF_RANGE,
                    "shell weights cannot be empty")
    return dict(sorted(w.items()))


def _array(key, value, S: int, K: int, single: bool = False) -> np.ndarray:
    """arrays.h:76-101: scalar, K x K matrix stacked over shells, or S x K x K cube"""
    a = np.asarray(value, dtype=np.float64)
    if single:  # the values of a configuration<float>
        a = a.astype(np.float32).astype(np.float64)
    if a.ndim == 0:
        return np.full((S, K, K), float(a))
    if a.ndim == 2:
        if a.shape != (K, K):
            raise _Fail(key, ParseErrorCode.CODE_BAD_VALUE,
                        f"matrix has invalid size {a.shape[0]}x{a.shape[1]} but expected {K}x{K}")
        return np.repeat(a[None, :, :], S, axis=0).copy()
    if a.ndim == 3:
        if a.shape != (S, K, K):
            raise _Fail(key, ParseErrorCode.CODE_BAD_VALUE,
                        f"array has invalid size {'x'.join(map(str, a.shape))} but expected "
                        f"{S}x{K}x{K}")
        return a.copy()
    raise _Fail(key, ParseErrorCode.CODE_BAD_VALUE, "Could not convert to 3D array")


def _per_problem(doc, key, n: int, mode: str):
    if key not in doc:
        return [None] * n
    v = doc[key]
    if mode == split:
        if not isinstance(v, (list, tuple)) or len(v) != n:
            raise _Fail(key, ParseErrorCode.CODE_TYPE_ERROR,
                        'You have specified "split" mode. In split mode array type parameter must '
                        "be specified per sublattice")
        return list(v)
    return [v]


def _parse(doc: dict) -> SqsConfiguration:
    for k in doc:
        if k not in KNOWN_KEYS:
            raise _Fail(k, ParseErrorCode.CODE_UNKNOWN, f'Unknown key "{k}"')
    prec = _mode(doc, "prec", (double, single), double)
    sp = prec == single  # configuration<float> (combined.h:283-295)
    it_mode = _mode(doc, "iteration_mode", (random, systematic), random)
    sl_mode = _mode(doc, "sublattice_mode", (interact, split), interact)
    if it_mode == systematic and sl_mode == split:
        raise _Fail("iteration_mode", ParseErrorCode.CODE_BAD_VALUE,
                    'It is not possible to do an exhaustive search on multiple sublattices in mode '
                    '"split"')
    st, sc = _parse_structure(doc, sp)
    subs = _parse_composition(doc, st, sl_mode)
    applied = _apply_composition(st, subs)
    nprob = 1 if sl_mode == interact else len(subs)
    structs = [applied] if sl_mode == interact else [applied.sliced(sl.sites) for sl in subs]
    geo_structs = [st] if sl_mode == interact else [st.sliced(sl.sites) for sl in subs]
    # shell radii (default "peak")
    if "shell_radii" in doc:
        spec = doc["shell_radii"]
        if sl_mode == split:
            if not isinstance(spec, (list, tuple)) or len(spec) != nprob:
                raise _Fail("shell_radii", ParseErrorCode.CODE_TYPE_ERROR,
                            "You have to specify the shell radii per sublattice")
            radii = [_radii(doc, s, g) for s, g in zip(spec, geo_structs)]
        else:
            radii = [_radii(doc, spec, geo_structs[0])]
    else:
        radii = [_radii(doc, "peak", g) for g in geo_structs]
    # iterations (combined.h:79-100)
    if it_mode == systematic:
        if len(subs) != 1:
            raise _Fail("iterations", ParseErrorCode.CODE_BAD_VALUE,
                        'In "systematic" iteration mode only one sublattice is allowed')
        total = num_permutations([applied.species[i] for i in subs[0].sites])
        if total > 2 ** 64 - 1:
            raise _Fail("iterations", ParseErrorCode.CODE_BAD_VALUE,
                        f"The number of permutations to test is {total}. I'm a pretty fast "
                        "program, but this is too much even for me ;=)")
        iterations = total
    else:
        iterations = int(doc.get("iterations", ITERATIONS_DEFAULT))
    # seed (combined.h:47-77)
    nseed = 1 if sl_mode == interact else len(subs)
    if "seed" in doc:
        if it_mode == systematic:
            raise _Fail("seed", ParseErrorCode.CODE_BAD_ARGUMENT,
                        'You cannot specify a seed if you have set the iteration mode to '
                        '"systematic".')
        s = doc["seed"]
        if isinstance(s, (list, tuple)):
            if len(s) != nseed:
                raise _Fail("seed", ParseErrorCode.CODE_BAD_VALUE,
                            f"The number of seeds provided ({len(s)}) does not match the number "
                            f"of sublattices ({nseed})")
            seed = [None if v is None else int(v) for v in s]
        else:
            seed = [int(s)] * nseed
    else:
        seed = [None] * nseed
    # shell weights
    wspec = doc.get("shell_weights")
    if sl_mode == split and isinstance(wspec, (list, tuple)):
        if len(wspec) != nprob:
            raise _Fail("shell_weights", ParseErrorCode.CODE_BAD_VALUE, "Cannot interpret value")
        weights = [_weights(w, len(r), sp) for w, r in zip(wspec, radii)]
    else:
        weights = [_weights(wspec, len(r), sp) for r in radii]
    # arrays
    prefs, pws, tgts = [], [], []
    p_spec = _per_problem(doc, "prefactors", nprob, sl_mode)
    w_spec = _per_problem(doc, "pair_weights", nprob, sl_mode)
    t_spec = _per_problem(doc, "target_objective", nprob, sl_mode)
    for q in range(nprob):
        S = len(weights[q])
        K = len(set(structs[q].species))
        if p_spec[q] is None:  # structure.h:260-265 on the applied / decomposed structure
            sm = structs[q].shell_matrix(radii[q])
            prefs.append(prefactors(sm, list(weights[q]), structs[q].species, sp))
        else:
            prefs.append(_array("prefactors", p_spec[q], S, K, sp))
        pws.append(_array("pair_weights", w_spec[q], S, K, sp) if w_spec[q] is not None
                   else np.repeat((np.ones((K, K)) - np.eye(K))[None], S, axis=0))
        tgts.append(_array("target_objective", t_spec[q], S, K, sp) if t_spec[q] is not None
                    else np.zeros((S, K, K)))
    # chunk size, threads, keep, max results
    chunk = int(doc.get("chunk_size", min(iterations, CHUNK_SIZE_DEFAULT)))
    if chunk > iterations:
        raise _Fail("chunk_size", ParseErrorCode.CODE_OUT_OF_RANGE,
                    f'"chunk_size" was set to {chunk} and iterations set to {iterations}')
    tc = doc.get("thread_config", os.cpu_count() or 1)
    threads = [int(t) for t in tc] if isinstance(tc, (list, tuple)) else [int(tc)]
    if any(t == 0 for t in threads):
        raise _Fail("thread_config", ParseErrorCode.CODE_BAD_VALUE,
                    "The number of threads cannot be set to 0")
    keep = int(doc.get("keep", 1))
    if keep == 0:
        raise _Fail("keep", ParseErrorCode.CODE_BAD_VALUE,
                    "The number of best kept structures must be greater than 0")
    mrpo = doc.get("max_results_per_objective")
    if mrpo is not None and int(mrpo) <= 0:
        raise _Fail("max_results_per_objective", ParseErrorCode.CODE_BAD_VALUE,
                    "The number of structures per objective must be a positive integer number")
    if any(s is not None for s in seed) and any(t > 1 for t in threads):
        raise _Fail("seed", ParseErrorCode.CODE_BAD_VALUE,
                    "A random seed is set but the thread configuration specifies more than 1 "
                    "thread; seeded runs require (currently) exactly 1 thread for reproducibility")
    cfg = SqsConfiguration(sl_mode, it_mode, seed, sc, subs, radii, weights, prefs, pws, tgts,
                           iterations, chunk, threads, keep,
                           None if mrpo is None else int(mrpo), prec)
    cfg._structure = st
    return cfg


def parse_config(config: Union[dict, str]) -> Union[SqsConfiguration, ParseError]:
    """python/sqsgenerator/_optimize.py:116-144 + io/config/combined.h:283-295"""
    if isinstance(config, str):
        config = json.loads(config)
    try:
        return _parse(dict(config))
    except _Fail as f:
        return f.err


# ---- run ----------------------------------------------------------------------------------

class _Arrays:
    """Flattened sqsg_config whose buffers stay alive for the duration of a call."""

    def __init__(self, cfg: SqsConfiguration):
        st = cfg.structure()
        self.lat = _f64(st.lattice)
        self.frac = _f64(st.frac_coords)
        self.species = np.ascontiguousarray(st.species, dtype=np.uint8)
        subs = cfg.composition
        self.site_off = np.cumsum([0] + [len(s.sites) for s in subs]).astype(np.uint32)
        self.sites = np.ascontiguousarray(np.concatenate([s.sites for s in subs]), dtype=np.uint32)
        self.comp_off = np.cumsum([0] + [len(s.composition) for s in subs]).astype(np.uint32)
        self.comp_z = np.ascontiguousarray([z for s in subs for z in s.composition], dtype=np.uint8)
        self.comp_n = np.ascontiguousarray([a for s in subs for a in s.composition.values()],
                                           dtype=np.uint32)
        self.radii_off = np.cumsum([0] + [len(r) for r in cfg.shell_radii]).astype(np.uint32)
        self.radii = _f64(np.concatenate([np.asarray(r, dtype=float) for r in cfg.shell_radii]))
        self.shell_off = np.cumsum([0] + [len(w) for w in cfg.shell_weights]).astype(np.uint32)
        self.shells = np.ascontiguousarray([s for w in cfg.shell_weights for s in w], np.uint32)
        self.weights = _f64([v for w in cfg.shell_weights for v in w.values()])
        self.K = np.ascontiguousarray([p.shape[1] for p in cfg.prefactors], dtype=np.uint32)
        self.pref = _f64(np.concatenate([p.ravel() for p in cfg.prefactors]))
        self.pw = _f64(np.concatenate([p.ravel() for p in cfg.pair_weights]))
        self.tgt = _f64(np.concatenate([p.ravel() for p in cfg.target_objective]))
        self.seeds = np.ascontiguousarray([s or 0 for s in cfg.seed], dtype=np.uint64)
        self.seed_present = np.ascontiguousarray([s is not None for s in cfg.seed], np.uint8)
        mrpo = cfg.max_results_per_objective
        self.c = L.Config(
            len(self.species), L.ptr(self.lat, L.f64p), L.ptr(self.frac, L.f64p),
            L.ptr(self.species, L.u8p),
            L.SUBLATTICE_SPLIT if cfg.sublattice_mode == split else L.SUBLATTICE_INTERACT,
            L.ITERATION_SYSTEMATIC if cfg.iteration_mode == systematic else L.ITERATION_RANDOM,
            len(subs), L.ptr(self.site_off, L.u32p), L.ptr(self.sites, L.u32p),
            L.ptr(self.comp_off, L.u32p), L.ptr(self.comp_z, L.u8p), L.ptr(self.comp_n, L.u32p),
            L.ptr(self.radii_off, L.u32p), L.ptr(self.radii, L.f64p),
            L.ptr(self.shell_off, L.u32p), L.ptr(self.shells, L.u32p),
            L.ptr(self.weights, L.f64p), L.ptr(self.K, L.u32p), L.ptr(self.pref, L.f64p),
            L.ptr(self.pw, L.f64p), L.ptr(self.tgt, L.f64p), cfg.iterations, cfg.chunk_size,
            L.ptr(self.seeds, L.u64p), L.ptr(self.seed_present, L.u8p), cfg.keep,
            -1 if mrpo is None else mrpo, 1 if cfg.prec == single else 0)


class Job:
    """Owning handle of an sqsg_job (one optimisation on the device)."""

    def __init__(self, cfg: SqsConfiguration):
        if isinstance(cfg, ParseError):
            raise ValueError(f"{cfg.key}: {cfg.msg}")
        self.config = cfg
        self._arrays = _Arrays(cfg)
        h = C.c_void_p()
        L.check(L.lib().sqsg_job_create(C.byref(self._arrays.c), C.byref(h)))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            L.lib().sqsg_job_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- introspection --
    def num_problems(self) -> int:
        n = C.c_uint32()
        L.check(L.lib().sqsg_job_num_problems(self.handle, C.byref(n)))
        return n.value

    def problem(self, p: int = 0) -> dict:
        info = L.ProblemInfo()
        L.check(L.lib().sqsg_job_problem_info(self.handle, p, C.byref(info)))
        N, K, S, P = info.num_sites, info.num_species, info.num_shells, info.num_pairs
        pairs = np.empty((P, 3), dtype=np.uint32)
        sp = np.empty(N, dtype=np.uint8)
        order = np.empty(N, dtype=np.uint32)
        pref, pw, tgt = (np.empty((S, K, K)) for _ in range(3))
        L.check(L.lib().sqsg_job_problem_arrays(self.handle, p, L.ptr(pairs, L.u32p),
                                                L.ptr(sp, L.u8p), L.ptr(order, L.u32p),
                                                L.ptr(pref, L.f64p), L.ptr(pw, L.f64p),
                                                L.ptr(tgt, L.f64p)))
        return dict(num_sites=N, num_species=K, num_shells=S, num_pairs=P, pairs=pairs,
                    species_packed=sp, sort_order=order, prefactors=pref, pair_weights=pw,
                    target=tgt, bounds=[(info.bounds[2 * b], info.bounds[2 * b + 1])
                                        for b in range(info.num_bounds)],
                    species_z=list(info.species_z[:K]), shell_ids=list(info.shell_ids[:S]))

    def score(self, species, p: int = 0, kernel: int = L.KERNEL_AUTO, bonds: bool = True):
        """Explicit candidates (B x N packed, sorted-site order) -> bonds (None unless
        requested), objective, sro"""
        info = self.problem(p)
        N, K, S = info["num_sites"], info["num_species"], info["num_shells"]
        sp = np.ascontiguousarray(species, dtype=np.uint8).reshape(-1, N)
        B = sp.shape[0]
        bd = np.empty((B, S, K, K), dtype=np.int32) if bonds else None
        obj = np.empty(B)
        sro = np.empty((B, S, K, K))
        L.check(L.lib().sqsg_score_batch(self.handle, p, L.ptr(sp, L.u8p), B,
                                         L.ptr(bd, L.i32p) if bonds else None,
                                         L.ptr(obj, L.f64p), L.ptr(sro, L.f64p), kernel))
        return bd, obj, sro

    def set_kernel(self, kernel: int) -> None:
        L.check(L.lib().sqsg_job_set_kernel(self.handle, kernel))

    def set_pruning(self, enable: bool = True) -> None:
        """early exit of passes that cannot reach the kept levels (identical results; needs
        non-negative pair weights)"""
        L.check(L.lib().sqsg_job_set_pruning(self.handle, int(enable)))

    def set_launch_size(self, n: int) -> None:
        L.check(L.lib().sqsg_job_set_launch_size(self.handle, n))

    def run(self, callback: Optional[Callable] = None) -> None:
        if callback is None:
            L.check(L.lib().sqsg_run(self.handle, L.PROGRESS_FN(), None))
            return
        job = self
        raised = []

        def trampoline(stats_p, _user):
            ctx = SqsCallbackContext(_stats_dict(stats_p.contents))
            try:
                callback(ctx)
            except BaseException as e:  # stop the run, re-raised once sqsg_run returns
                raised.append(e)
                return 1
            return 1 if ctx._stop else 0

        fn = L.PROGRESS_FN(trampoline)
        L.check(L.lib().sqsg_run(job.handle, fn, None))
        if raised:
            raise raised[0]

    def run_noalloc(self) -> None:
        """the same run without device allocations (sqsg_run_noalloc): the job's inputs must
        be resident already, i.e. after a first run()"""
        L.check(L.lib().sqsg_run_noalloc(self.handle, L.PROGRESS_FN(), None))

    def cancel(self) -> None:
        L.lib().sqsg_cancel(self.handle)

    def statistics(self) -> dict:
        s = L.Statistics()
        L.check(L.lib().sqsg_job_statistics(self.handle, C.byref(s)))
        return _stats_dict(s)

    def timing(self) -> tuple[float, int, int]:
        ms, n, c = C.c_double(), C.c_uint64(), C.c_uint64()
        L.check(L.lib().sqsg_job_timing(self.handle, C.byref(ms), C.byref(n), C.byref(c)))
        return ms.value, n.value, c.value

    def results(self) -> "SqsResultPack":
        return SqsResultPack._from_job(self)

    def kernel_shape(self) -> str:
        """launch shape of the kernel a run uses, e.g. "dense<8,2,2,3>" """
        buf = C.create_string_buffer(128)
        L.check(L.lib().sqsg_job_kernel_shape(self.handle, buf, len(buf)))
        return buf.value.decode()

    def autotune(self) -> tuple[int, float]:
        """times the kernel variants and the dense launch parameters on probes of the job and
        keeps the fastest (clears results); returns the variant and its probe rate"""
        k, r = C.c_int(), C.c_double()
        L.check(L.lib().sqsg_autotune(self.handle, C.byref(k), C.byref(r)))
        return k.value, r.value

    def set_dense_launch(self, shape: int = 0, passes: int = 0, launch: int = 0) -> None:
        """dense search launch shape (0: default, 1..: alternatives of the tile bound),
        candidate passes per wavefront round (0: as many as fit) and candidates per launch
        (0: default); results are identical for every setting"""
        L.check(L.lib().sqsg_job_set_dense_launch(self.handle, shape, passes, launch))

    def dense_launch(self) -> dict:
        n, s, p, l = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        L.check(L.lib().sqsg_job_dense_launch(self.handle, C.byref(n), C.byref(s), C.byref(p),
                                              C.byref(l)))
        return dict(num_shapes=n.value, shape=s.value, passes=p.value, launch=l.value)


def batch_run(jobs: list["Job"], use_graph: bool = True) -> dict:
    """Runs every job to completion in shared persistent launches (the batched form)."""
    arr = (C.c_void_p * len(jobs))(*[j.handle for j in jobs])
    rep = L.BatchReport()
    L.check(L.lib().sqsg_batch_run(arr, len(jobs), int(use_graph), C.byref(rep)))
    return dict(kernel_ms=rep.kernel_ms, candidates=rep.candidates, launches=rep.launches,
                work_items=rep.work_items)


class Server:
    """Continuous batching: submitted jobs are admitted by device-memory footprint between
    persistent steps and advance together with the jobs already resident."""

    def __init__(self, hbm_budget_bytes: int = 0, max_active: int = 64,
                 slice_candidates: int = 0):
        opt = L.ServerOptions(hbm_budget_bytes, max_active, slice_candidates)
        h = C.c_void_p()
        L.check(L.lib().sqsg_server_create(C.byref(opt), C.byref(h)))
        self.handle = h
        self._jobs = {}

    def submit(self, job: "Job") -> int:
        t = C.c_uint64()
        L.check(L.lib().sqsg_server_submit(self.handle, job.handle, C.byref(t)))
        self._jobs[t.value] = job
        return t.value

    def step(self) -> tuple[int, int]:
        a, q = C.c_int(), C.c_int()
        L.check(L.lib().sqsg_server_step(self.handle, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, ticket: int) -> dict:
        d, ts, ta, td = C.c_int(), C.c_double(), C.c_double(), C.c_double()
        L.check(L.lib().sqsg_server_poll(self.handle, ticket, C.byref(d), C.byref(ts),
                                         C.byref(ta), C.byref(td)))
        return dict(done=bool(d.value), t_submit=ts.value, t_admit=ta.value, t_done=td.value)

    def drain(self) -> None:
        while True:
            a, q = self.step()
            if a == 0 and q == 0:
                return

    def close(self) -> None:
        if self.handle:
            L.lib().sqsg_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stats_dict(s) -> dict:
    return dict(finished=s.finished, working=s.working, best_rank=s.best_rank,
                best_objective=s.best_objective,
                timings=dict(total=s.timing_total_ns, chunk_setup=s.timing_chunk_setup_ns,
                             loop=s.timing_loop_ns, comm=s.timing_comm_ns))


class SqsCallbackContext:
    """types.h:222-232"""

    def __init__(self, statistics: dict):
        self.statistics = statistics
        self._stop = False

    def stop(self) -> None:
        self._stop = True


class SqsResult:
    """One reported configuration (core/results.h:142-283).  Species are in the original
    site order as atomic numbers; SRO and rank are fetched from the job on first use."""

    def __init__(self, pack: "SqsResultPack", index: int, objective: float, species):
        self._pack = pack
        self.index = index
        self.objective = objective
        self.species = [int(z) for z in species]
        self._sro = None
        self._sub = None
        self._rank = None
        self._cand = None

    def _fetch(self):
        job = self._pack._job
        nprob = len(self._pack._problems)
        sub = np.empty(nprob)
        sro = np.empty(self._pack._cube_total)
        rank = C.create_string_buffer(1 << 16)
        cand = C.c_uint64()
        L.check(L.lib().sqsg_result_get(job.handle, self.index, None, None, L.ptr(sub, L.f64p),
                                        L.ptr(sro, L.f64p), rank, len(rank), C.byref(cand)))
        cubes, off = [], 0
        for p in self._pack._problems:
            n = p["num_shells"] * p["num_species"] ** 2
            cubes.append(sro[off:off + n].reshape(p["num_shells"], p["num_species"],
                                                  p["num_species"]))
            off += n
        self._sro, self._sub, self._rank, self._cand = cubes, sub, rank.value.decode(), cand.value

    @property
    def candidate(self) -> int:
        if self._cand is None:
            self._fetch()
        return self._cand

    def sro(self, *args):
        if self._pack.config.sublattice_mode == split:
            raise ValueError("split-mode results expose SRO per sublattice (.sublattices)")
        if self._sro is None:
            self._fetch()
        return self._sro[0].copy() if not args else self.parameter(*args)

    def parameter(self, shell: int, i=None, j=None):
        if self._sro is None:
            self._fetch()
        info = self._pack._problems[0]
        s = info["shell_ids"].index(shell)
        if i is None:
            return self._sro[0][s].copy()
        a, b = info["species_z"].index(to_z(i)), info["species_z"].index(to_z(j))
        return float(self._sro[0][s, a, b])

    def rank(self) -> str:
        if self._rank is None:
            self._fetch()
        return self._rank

    def structure(self) -> Structure:
        return self._pack.config.structure().with_species(self.species)

    @property
    def sublattices(self) -> list["SqsSublatticeResult"]:
        if self._sro is None:
            self._fetch()
        ranks = self._rank.split(",")
        return [SqsSublatticeResult(float(self._sub[q]), self._sro[q],
                                    ranks[q] if q < len(ranks) else "", info)
                for q, info in enumerate(self._pack._problems)]


@dataclass
class SqsSublatticeResult:
    objective: float
    _sro: np.ndarray
    _rank: str
    _info: dict

    def sro(self):
        return self._sro.copy()

    def rank(self) -> str:
        return self._rank


class SqsResultPack:
    """core/results.h:336-389: (objective, [results]) in ascending objective order.

    The results stay in the job (host memory of the C++ layer); levels can hold millions of
    ties, so per-result objects are built only when a level is iterated, and the bulk
    accessors return numpy arrays."""

    def __init__(self, job: "Job"):
        self._job = job
        self.config = job.config
        self.statistics = job.statistics()
        self._problems = [job.problem(q) for q in range(job.num_problems())]
        self._cube_total = sum(p["num_shells"] * p["num_species"] ** 2 for p in self._problems)
        nr, nl = C.c_uint64(), C.c_uint64()
        L.check(L.lib().sqsg_result_count(job.handle, C.byref(nr), C.byref(nl)))
        self._n = nr.value
        self.level_objectives = np.empty(nl.value)
        self.level_start = np.empty(nl.value + 1, dtype=np.uint64)
        L.check(L.lib().sqsg_result_levels(job.handle, L.ptr(self.level_objectives, L.f64p),
                                           L.ptr(self.level_start, L.u64p)))

    @classmethod
    def _from_job(cls, job: "Job") -> "SqsResultPack":
        return cls(job)

    def arrays(self, first: int = 0, count: Optional[int] = None, species: bool = True,
               ranks: bool = True) -> dict:
        """objective, candidate, original-order species (Z) and uint64 rank arrays"""
        count = self._n - first if count is None else count
        N = len(self.config.structure())
        obj = np.empty(count)
        cand = np.empty(count, dtype=np.uint64)
        sp = np.empty((count, N), dtype=np.uint8) if species else None
        rk = np.empty(count, dtype=np.uint64) if ranks else None
        L.check(L.lib().sqsg_result_bulk(self._job.handle, first, count, L.ptr(obj, L.f64p),
                                         L.ptr(cand, L.u64p),
                                         L.ptr(sp, L.u8p) if species else None,
                                         L.ptr(rk, L.u64p) if ranks else None))
        return dict(objective=obj, candidate=cand, species=sp, rank=rk)

    def species_packed(self, r: int) -> np.ndarray:
        """internal form of result r: packed species indices in sorted-site order,
        concatenated over the problems (the configuration the kernels scored)"""
        out = np.empty(sum(p["num_sites"] for p in self._problems), dtype=np.uint8)
        L.check(L.lib().sqsg_result_species_packed(self._job.handle, r, L.ptr(out, L.u8p)))
        return out

    def sro_array(self, first: int = 0, count: Optional[int] = None) -> np.ndarray:
        count = self._n - first if count is None else count
        out = np.empty((count, self._cube_total))
        L.check(L.lib().sqsg_result_sro_bulk(self._job.handle, first, count, L.ptr(out, L.f64p)))
        return out

    def _level(self, l: int) -> list[SqsResult]:
        a, b = int(self.level_start[l]), int(self.level_start[l + 1])
        arr = self.arrays(a, b - a, species=True, ranks=False)
        return [SqsResult(self, a + k, float(arr["objective"][k]), arr["species"][k])
                for k in range(b - a)]

    def __iter__(self):
        for l in range(len(self.level_objectives)):
            yield float(self.level_objectives[l]), self._level(l)

    def __len__(self) -> int:
        return len(self.level_objectives)

    def num_results(self) -> int:
        return self._n

    def best(self) -> SqsResult:
        return self._level(0)[0]

    def _result_docs(self) -> list:
        """the flattened results of the wire format (io/json.h:221-262), level order"""
        from . import pack_io as W
        if self._n == 0:
            return []
        arr = self.arrays(0, self._n, species=True, ranks=False)
        sro = self.sro_array(0, self._n)
        sub = np.empty((self._n, len(self._problems)))
        if self.config.sublattice_mode == split:
            for i in range(self._n):  # per-sublattice objectives
                r = SqsResult(self, i, float(arr["objective"][i]), arr["species"][i])
                r._fetch()
                sub[i] = r._sub
        shapes = [(p["num_shells"], p["num_species"]) for p in self._problems]
        docs = []
        for i in range(self._n):
            cubes, off = [], 0
            for S_, K_ in shapes:
                cubes.append(sro[i, off:off + S_ * K_ * K_].reshape(S_, K_, K_))
                off += S_ * K_ * K_
            sp = arr["species"][i]
            if self.config.sublattice_mode == split:
                subs = [W.result_doc(sub[i, q], [int(sp[j]) for j in sl.sites], cubes[q])
                        for q, sl in enumerate(self.config.composition)]
                docs.append(W._obj(objective=float(arr["objective"][i]), sublattices=subs))
            else:
                docs.append(W.result_doc(arr["objective"][i], sp, cubes[0]))
        return docs

    def bytes(self) -> bytes:
        """msgpack wire form of the pack (sqsgen.cpp:67-73 to_bytes, io/json.h:264-276)"""
        from . import pack_io as W
        return W.dumps(W.pack_doc(self.config, self.statistics, self._result_docs()))

    @staticmethod
    def from_bytes(data: bytes) -> "StoredResultPack":
        return StoredResultPack.from_bytes(data)


class StoredResult:
    """A result read from the wire format: objective, original-order species (Z) and the SRO
    cube(s) it carries; the rank is recomputed from the species (results.h:163-169)."""

    def __init__(self, pack: "StoredResultPack", objective: float, species, sro, subs=None):
        self._pack = pack
        self.objective = objective
        self.species = [int(z) for z in species]
        self._sro = sro
        self._subs = subs

    def sro(self, *args):
        if self._pack.config.sublattice_mode == split:
            raise ValueError("split-mode results expose SRO per sublattice (.sublattices)")
        return self._sro.copy() if not args else self.parameter(*args)

    def parameter(self, shell: int, i=None, j=None):
        shells = sorted(self._pack.config.shell_weights[0])
        s = shells.index(shell)
        if i is None:
            return self._sro[s].copy()
        zs = sorted(set(self.species))
        return float(self._sro[s, zs.index(to_z(i)), zs.index(to_z(j))])

    def rank(self) -> str:
        from . import pack_io as W
        if self._pack.config.sublattice_mode == split:
            return ",".join(r.rank() for r in self.sublattices)
        return W.species_rank(self.species)

    def structure(self) -> Structure:
        return self._pack.config.structure().with_species(self.species)

    @property
    def sublattices(self) -> list["SqsSublatticeResult"]:
        from . import pack_io as W
        if self._subs is None:
            return []
        return [SqsSublatticeResult(float(o), cube, W.species_rank(sp), {})
                for o, sp, cube in self._subs]


class StoredResultPack:
    """A result pack read from the wire format (load_result_pack, sqsgen.cpp:38-56): the same
    iteration and accessors as SqsResultPack, held in host memory."""

    def __init__(self, config: SqsConfiguration, statistics: dict, results: list):
        from . import pack_io as W
        self.config = config
        self.statistics = statistics
        is_split = config.sublattice_mode == split
        objs, levels = W.levels_from_results(results, is_split)
        self.level_objectives = np.asarray(objs, dtype=np.float64)
        self._levels = []
        for docs in levels:
            lv = []
            for r in docs:
                if is_split:
                    subs = [(float(sl["objective"]), [int(z) for z in sl["species"]],
                             W.cube_from_doc(sl["sro"])) for sl in r["sublattices"]]
                    species = list(config.structure().species)
                    for (_, sp, _c), sl in zip(subs, config.composition):
                        for site, z in zip(sl.sites, sp):
                            species[site] = z
                    lv.append(StoredResult(self, float(r["objective"]), species, None, subs))
                else:
                    lv.append(StoredResult(self, float(r["objective"]), r["species"],
                                           W.cube_from_doc(r["sro"])))
            self._levels.append(lv)
        self._n = sum(len(lv) for lv in self._levels)

    @classmethod
    def from_bytes(cls, data: bytes) -> "StoredResultPack":
        from . import pack_io as W
        doc = W.loads(bytes(data))
        try:
            cfg = W.config_from_doc(doc["config"])
            stats = W.statistics_from_doc(doc["statistics"])
            results = doc["results"]
        except (KeyError, TypeError, ValueError) as e:
            raise ValueError(f"Invalid result pack: {e}") from e
        if cfg.sublattice_mode not in (interact, split):
            raise ValueError("Invalid result pack - invalid sublattice_mode")
        return cls(cfg, stats, results)

    def bytes(self) -> bytes:
        from . import pack_io as W
        docs = []
        for lv in self._levels:
            for r in lv:
                if r._subs is not None:
                    docs.append(W._obj(objective=r.objective, sublattices=[
                        W.result_doc(o, sp, cube) for o, sp, cube in r._subs]))
                else:
                    docs.append(W.result_doc(r.objective, r.species, r._sro))
        return W.dumps(W.pack_doc(self.config, self.statistics, docs))

    def __iter__(self):
        for l, lv in enumerate(self._levels):
            yield float(self.level_objectives[l]), list(lv)

    def __len__(self) -> int:
        return len(self._levels)

    def num_results(self) -> int:
        return self._n

    def best(self) -> StoredResult:
        return self._levels[0][0]


def load_result_pack(data, prec: str = double) -> StoredResultPack:
    """python/extension/sqsgen.cpp:38-56,528-531: a msgpack result pack from a file path or
    from the bytes themselves"""
    from . import pack_io as W
    if prec not in (double, single):
        raise ValueError(f"unknown precision {prec!r}")
    pack = StoredResultPack.from_bytes(W.read_pack_input(data))
    # the wire form carries no precision (io/json.h:177-195): the caller names it
    pack.config.prec = prec
    return pack


def optimize(config: Union[dict, SqsConfiguration, str], level: Any = None,
             callback: Optional[Callable[[SqsCallbackContext], None]] = None) -> SqsResultPack:
    """python/sqsgenerator/_optimize.py:147-168 -> run_optimization (sqs.h:625-634)"""
    cfg = config if isinstance(config, SqsConfiguration) else parse_config(config)
    if isinstance(cfg, ParseError):
        raise ValueError(f"{cfg.key}: {cfg.msg}")
    job = Job(cfg)
    job.run(callback)
    return job.results()
