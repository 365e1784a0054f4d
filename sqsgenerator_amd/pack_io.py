"""The reference's binary wire format for structures, configurations and result packs.

sqsgenerator serialises these objects as msgpack of the JSON document nlohmann::json builds
for them (io/json.h:33-51,120-287; python/extension/sqsgen.cpp:29-77):

* a result pack is ``{"config", "results", "statistics"}`` with the results flattened in
  ascending objective order (io/json.h:264-276);
* an interact-mode result is ``{"objective", "species", "sro"}``: species as atomic numbers
  in the original site order (sqs.h:94-122), ``sro`` as ``{"shape": [S, K, K], "data": [...]}``
  with the cube's elements in Eigen's column-major order, element (s, a, b) at
  ``s + S * (a + K * b)`` (io/json.h:33-51);
* a split-mode result is ``{"objective", "sublattices": [interact-mode results]}``;
* maps whose keys are not strings (composition Z -> amount, shell weights, timings) are
  arrays of ``[key, value]`` pairs in ascending key order; nlohmann objects are std::maps, so
  object keys are written in ascending order too.

The helpers here build and read those documents; msgpack does the byte encoding.  Readers of
either side accept float32 or float64 numbers, so documents written here load in the
reference (``load_result_pack``, ``SqsResultPack.from_bytes``) and the other way round.
"""
from __future__ import annotations

import os
from typing import Any, Optional, Union

import msgpack
import numpy as np

# Timing enum values (types.h:182-188) order the statistics' timing pairs
TIMINGS = ("total", "chunk_setup", "loop", "comm")


def _obj(**kw) -> dict:
    """a JSON object with its keys in the order nlohmann writes them (std::map: ascending)"""
    return {k: kw[k] for k in sorted(kw)}


def _pairs(d: dict) -> list:
    return [[k, d[k]] for k in sorted(d)]


def _unpairs(v) -> dict:
    if isinstance(v, dict):  # tolerate the object form
        return {int(k): x for k, x in v.items()}
    return {k: x for k, x in v}


def cube_doc(cube) -> dict:
    """binary_adapter<T, 3>::save (io/json.h:33-40): shape plus column-major data"""
    c = np.asarray(cube, dtype=np.float64)
    if c.ndim != 3:
        raise ValueError("an SRO cube has three dimensions [S][K][K]")
    return _obj(shape=[int(x) for x in c.shape], data=c.reshape(-1, order="F").tolist())


def cube_from_doc(doc) -> np.ndarray:
    """binary_adapter<T, 3>::load (io/json.h:42-50)"""
    shape = [int(x) for x in doc["shape"]]
    data = np.asarray(doc["data"], dtype=np.float64)
    if len(shape) != 3 or data.size != int(np.prod(shape)):
        raise ValueError("invalid SRO cube: shape and data size disagree")
    return data.reshape(shape, order="F")


def dumps(doc: Any) -> bytes:
    return msgpack.packb(doc, use_bin_type=True)


def loads(data: bytes) -> Any:
    return msgpack.unpackb(data, raw=False, strict_map_key=False)


# ---- structures and configurations ---------------------------------------------------------

def structure_doc(st) -> dict:
    """adl_serializer<core::structure<T>> (io/json.h:139-157)"""
    return _obj(lattice=np.asarray(st.lattice, dtype=np.float64).tolist(),
                species=[int(z) for z in st.species],
                frac_coords=np.asarray(st.frac_coords, dtype=np.float64).tolist(),
                pbc=[bool(x) for x in st.pbc])


def structure_from_doc(doc):
    from .core import Structure
    return Structure(doc["lattice"], doc["frac_coords"], doc["species"],
                     tuple(doc.get("pbc", (True, True, True))))


def config_doc(cfg) -> dict:
    """adl_serializer<core::configuration<T>> (io/json.h:177-195)"""
    sc = cfg.structure_config
    structure = _obj(lattice=[list(map(float, r)) for r in sc["lattice"]],
                     species=[int(z) for z in sc["species"]],
                     coords=[list(map(float, r)) for r in sc["coords"]],
                     supercell=[int(x) for x in sc["supercell"]])
    composition = [_obj(sites=sorted(int(i) for i in sl.sites),
                        composition=_pairs({int(z): int(a) for z, a in sl.composition.items()}))
                   for sl in cfg.composition]
    cubes = lambda arrs: [np.asarray(a, dtype=np.float64).tolist() for a in arrs]
    return _obj(sublattice_mode=cfg.sublattice_mode, iteration_mode=cfg.iteration_mode,
                seed=[None if s is None else int(s) for s in cfg.seed], structure=structure,
                composition=composition,
                shell_radii=[[float(r) for r in radii] for radii in cfg.shell_radii],
                shell_weights=[_pairs({int(k): float(v) for k, v in w.items()})
                               for w in cfg.shell_weights],
                prefactors=cubes(cfg.prefactors), pair_weights=cubes(cfg.pair_weights),
                target_objective=cubes(cfg.target_objective),
                iterations=None if cfg.iterations is None else int(cfg.iterations),
                chunk_size=int(cfg.chunk_size),
                thread_config=[int(t) for t in cfg.thread_config], keep=int(cfg.keep),
                max_results_per_objective=None if cfg.max_results_per_objective is None
                else int(cfg.max_results_per_objective))


def config_from_doc(doc):
    """configuration<T> from_json (io/json.h:197-213): every key is required"""
    from .core import SqsConfiguration, Structure, Sublattice
    s = doc["structure"]
    sc = {"lattice": [list(map(float, r)) for r in s["lattice"]],
          "coords": [list(map(float, r)) for r in s["coords"]],
          "species": [int(z) for z in s["species"]],
          "supercell": [int(x) for x in s["supercell"]]}
    subs = [Sublattice([int(i) for i in sl["sites"]],
                       {int(z): int(a) for z, a in sorted(_unpairs(sl["composition"]).items())})
            for sl in doc["composition"]]
    to_cubes = lambda v: [np.asarray(c, dtype=np.float64) for c in v]
    cfg = SqsConfiguration(
        doc["sublattice_mode"], doc["iteration_mode"],
        [None if x is None else int(x) for x in doc["seed"]], sc, subs,
        [[float(r) for r in radii] for radii in doc["shell_radii"]],
        [{int(k): float(v) for k, v in sorted(_unpairs(w).items())} for w in doc["shell_weights"]],
        to_cubes(doc["prefactors"]), to_cubes(doc["pair_weights"]),
        to_cubes(doc["target_objective"]),
        None if doc["iterations"] is None else int(doc["iterations"]), int(doc["chunk_size"]),
        [int(t) for t in doc["thread_config"]], int(doc["keep"]),
        None if doc["max_results_per_objective"] is None
        else int(doc["max_results_per_objective"]))
    cfg._structure = Structure(sc["lattice"], sc["coords"], sc["species"]).supercell(
        *sc["supercell"])
    return cfg


def statistics_doc(stats: dict) -> dict:
    """adl_serializer<sqs_statistics_data<T>> (io/json.h:119-135); the timings map is keyed by
    the Timing enum, so it is an array of [name, nanoseconds] pairs in enum order"""
    t = stats.get("timings", {})
    rank = stats.get("best_rank", 0)
    return _obj(best_objective=float(stats.get("best_objective", float("inf"))),
                best_rank=int(rank) if str(rank).isdigit() else 0,
                finished=int(stats.get("finished", 0)), working=int(stats.get("working", 0)),
                timings=[[k, int(t.get(k, 0))] for k in TIMINGS])


def statistics_from_doc(doc) -> dict:
    t = doc.get("timings", [])
    timings = dict(t) if isinstance(t, list) else dict(t)
    return dict(best_objective=float(doc["best_objective"]), best_rank=int(doc["best_rank"]),
                finished=int(doc["finished"]), working=int(doc["working"]),
                timings={k: int(timings.get(k, 0)) for k in TIMINGS})


# ---- result packs --------------------------------------------------------------------------

def result_doc(objective: float, species, sro) -> dict:
    return _obj(objective=float(objective), species=[int(z) for z in species],
                sro=cube_doc(sro))


def pack_doc(config, statistics: dict, results: list) -> dict:
    """adl_serializer<sqs_result_pack<T, Mode>> (io/json.h:264-276); `results` are already
    flattened documents in ascending objective order"""
    return _obj(config=config_doc(config), results=results,
                statistics=statistics_doc(statistics))


def read_pack_input(data: Union[bytes, bytearray, str, os.PathLike]) -> bytes:
    """load_result_pack (sqsgen.cpp:38-56): a path to an existing file, or the bytes
    themselves"""
    if isinstance(data, (bytes, bytearray)):
        if len(data) < 4096:
            try:
                if os.path.exists(os.fsdecode(bytes(data))):
                    with open(os.fsdecode(bytes(data)), "rb") as f:
                        return f.read()
            except (ValueError, OSError):
                pass
        return bytes(data)
    path = os.fspath(data)
    with open(path, "rb") as f:
        return f.read()


def levels_from_results(results: list, split: bool) -> tuple[list, list]:
    """groups flattened result documents into ascending objective levels (the pack's
    sqs_result_collection insert, io/json.h:279-281; objectives compare exactly)"""
    order: dict[float, list] = {}
    for r in results:
        order.setdefault(float(r["objective"]), []).append(r)
    objs = sorted(order)
    return objs, [order[o] for o in objs]


def species_rank(species) -> str:
    """sqs_result_wrapper::rank (results.h:163-169): rank of the configuration with species
    mapped to 0..K-1 in ascending atomic number"""
    from .core import rank_permutation
    zs = sorted(set(int(z) for z in species))
    return str(rank_permutation([zs.index(int(z)) for z in species]))
