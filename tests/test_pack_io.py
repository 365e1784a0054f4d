"""The reference's msgpack wire format for structures, configurations and result packs
(io/json.h:33-51,120-287; python/extension/sqsgen.cpp:29-77), on the CPU.

The reference writes these documents with nlohmann::json::to_msgpack; no packed file ships
with it, so the known-answer document below is written the way nlohmann lays it out: object
keys ascending, maps with non-string keys as [key, value] pairs, SRO cubes column-major, and
numbers that fit a float32 exactly written as float32 (nlohmann's compact floats).
"""
import math

import msgpack
import numpy as np
import pytest

import workloads as W
import sqsgenerator_amd as S
from sqsgenerator_amd import pack_io as P
from sqsgenerator_amd.core import StoredResultPack


def test_cube_is_column_major():
    """binary_adapter<T, 3>::save writes Eigen's storage order: (s, a, b) at s + S(a + K b)"""
    S_, K = 3, 4
    cube = np.arange(S_ * K * K, dtype=np.float64).reshape(S_, K, K) * 0.25 + 1.0
    doc = P.cube_doc(cube)
    assert list(doc) == ["data", "shape"]
    assert doc["shape"] == [S_, K, K]
    for s in range(S_):
        for a in range(K):
            for b in range(K):
                assert doc["data"][s + S_ * (a + K * b)] == cube[s, a, b]
    assert np.array_equal(P.cube_from_doc(doc), cube)


def _config():
    d = W.plain_mc_32(iterations=1000)
    d["chunk_size"] = 1000
    cfg = S.parse_config(d)
    assert not isinstance(cfg, S.ParseError)
    return cfg


def test_configuration_document_layout():
    """configuration<T> to_json (io/json.h:177-195): every key the reader's from_json asks for,
    ascending key order, maps with integer keys as pair arrays"""
    cfg = _config()
    doc = P.loads(cfg.bytes())
    keys = ["chunk_size", "composition", "iteration_mode", "iterations", "keep",
            "max_results_per_objective", "pair_weights", "prefactors", "seed", "shell_radii",
            "shell_weights", "structure", "sublattice_mode", "target_objective",
            "thread_config"]
    assert list(doc) == keys
    assert list(doc["structure"]) == ["coords", "lattice", "species", "supercell"]
    sl = doc["composition"][0]
    assert list(sl) == ["composition", "sites"]
    assert sl["composition"] == [[13, 16], [22, 16]]           # Al:16, Ti:16 as [Z, amount]
    assert doc["shell_weights"][0] == [[1, 1.0], [2, 0.5], [3, pytest.approx(1 / 3)]]
    assert doc["sublattice_mode"] == "interact" and doc["iteration_mode"] == "random"
    assert np.asarray(doc["prefactors"][0]).shape == (3, 2, 2)


def test_configuration_round_trip():
    cfg = _config()
    back = S.SqsConfiguration.from_bytes(cfg.bytes())
    for f in ("sublattice_mode", "iteration_mode", "seed", "structure_config", "shell_radii",
              "shell_weights", "iterations", "chunk_size", "thread_config", "keep",
              "max_results_per_objective"):
        assert getattr(back, f) == getattr(cfg, f), f
    for f in ("prefactors", "pair_weights", "target_objective"):
        for x, y in zip(getattr(back, f), getattr(cfg, f)):
            assert np.array_equal(x, y), f
    assert [(sl.sites, sl.composition) for sl in back.composition] == \
           [(sl.sites, sl.composition) for sl in cfg.composition]
    assert back.structure().species == cfg.structure().species
    assert np.array_equal(back.structure().frac_coords, cfg.structure().frac_coords)


def test_structure_round_trip():
    st = _config().structure()
    doc = P.loads(st.bytes())
    assert list(doc) == ["frac_coords", "lattice", "pbc", "species"]
    back = S.Structure.from_bytes(st.bytes())
    assert back.species == st.species and np.array_equal(back.lattice, st.lattice)
    assert np.array_equal(back.frac_coords, st.frac_coords) and back.pbc == st.pbc


def _nlohmann_pack(cfg):
    """a two-level pack written as the reference's writer lays it out"""
    doc = P.loads(cfg.bytes())
    N = len(cfg.structure())
    base = [13] * 16 + [22] * 16
    r1 = list(reversed(base))
    r2 = base[1:] + base[:1]
    cube = np.linspace(-1.0, 1.0, 12).reshape(3, 2, 2)

    def res(obj, sp, c):
        return {"objective": obj, "species": sp,
                "sro": {"data": c.reshape(-1, order="F").tolist(), "shape": [3, 2, 2]}}
    results = [res(0.5, base, cube), res(0.5, r1, cube * 0.5), res(1.25, r2, cube * 2.0)]
    stats = {"best_objective": 0.5, "best_rank": 7, "finished": 1000, "timings":
             [["total", 10], ["chunk_setup", 1], ["loop", 8], ["comm", 0]], "working": 0}
    assert N == len(base)
    return {"config": doc, "results": results, "statistics": stats}, (base, r1, r2, cube)


def _pack_bytes(doc):
    # nlohmann writes numbers that fit a float32 exactly as float32
    return msgpack.packb(doc, use_bin_type=True, use_single_float=True)


def test_load_reference_layout_known_answer(tmp_path):
    cfg = _config()
    doc, (base, r1, r2, cube) = _nlohmann_pack(cfg)
    data = _pack_bytes(doc)
    for src in (data, str(tmp_path / "pack.mpack")):
        if isinstance(src, str):
            with open(src, "wb") as f:
                f.write(data)
        pack = S.load_result_pack(src)
        assert isinstance(pack, StoredResultPack)
        assert pack.num_results() == 3 and len(pack) == 2
        levels = list(pack)
        assert [o for o, _ in levels] == [0.5, 1.25]
        assert sorted(tuple(r.species) for r in levels[0][1]) == sorted([tuple(base), tuple(r1)])
        top = [r for r in levels[0][1] if r.species == base][0]
        assert np.allclose(top.sro(), cube)          # float32-encoded data, float64 cube
        assert top.sro()[1, 0, 1] == pytest.approx(cube[1, 0, 1])
        # ranks are recomputed from the species (results.h:163-169): the sorted configuration
        # is rank 1, its reverse the last of C(32, 16)
        assert top.rank() == "1"
        last = [r for r in levels[0][1] if r.species == r1][0]
        assert last.rank() == str(math.comb(32, 16))
        assert pack.statistics["finished"] == 1000 and pack.statistics["best_rank"] == 7
        assert pack.statistics["timings"]["loop"] == 8
        assert pack.config.iterations == cfg.iterations


def test_stored_pack_round_trip():
    cfg = _config()
    doc, _ = _nlohmann_pack(cfg)
    pack = S.load_result_pack(_pack_bytes(doc))
    again = S.load_result_pack(pack.bytes())
    a = [(o, sorted((tuple(r.species), tuple(r.sro().ravel())) for r in rs)) for o, rs in pack]
    b = [(o, sorted((tuple(r.species), tuple(r.sro().ravel())) for r in rs)) for o, rs in again]
    assert a == b
    raw = P.loads(pack.bytes())
    assert list(raw) == ["config", "results", "statistics"]
    assert list(raw["results"][0]) == ["objective", "species", "sro"]
    assert raw["statistics"]["timings"] == [["total", 10], ["chunk_setup", 1], ["loop", 8],
                                            ["comm", 0]]


def test_split_result_layout():
    d = W.rocksalt_split_216(iterations=1000)
    d["chunk_size"] = 1000
    cfg = S.parse_config(d)
    assert not isinstance(cfg, S.ParseError)
    doc = P.loads(cfg.bytes())
    st = cfg.structure()
    subs = []
    for q, sl in enumerate(cfg.composition):
        K = len(sl.composition)
        Sq = len(cfg.shell_weights[q])
        sp = [z for z, a in sorted(sl.composition.items()) for _ in range(a)]
        subs.append({"objective": 0.25 * (q + 1), "species": sp,
                     "sro": P.cube_doc(np.full((Sq, K, K), 0.125 * (q + 1)))})
    pack_doc = {"config": doc, "results": [{"objective": 0.75, "sublattices": subs}],
                "statistics": P.statistics_doc({})}
    pack = S.load_result_pack(msgpack.packb(pack_doc, use_bin_type=True))
    (obj, rs), = list(pack)
    assert obj == 0.75 and len(rs) == 1
    r = rs[0]
    assert [s.objective for s in r.sublattices] == [0.25, 0.5]
    assert r.sublattices[1].sro()[0, 0, 0] == 0.25
    # the sublattice species land on their sites of the structure
    for q, sl in enumerate(cfg.composition):
        assert [r.species[i] for i in sl.sites] == subs[q]["species"]
    assert len(r.species) == len(st)
    with pytest.raises(ValueError):
        r.sro()
    back = P.loads(pack.bytes())
    assert list(back["results"][0]) == ["objective", "sublattices"]


def test_invalid_pack_raises():
    cfg = _config()
    doc, _ = _nlohmann_pack(cfg)
    del doc["config"]["keep"]
    with pytest.raises(ValueError):
        S.load_result_pack(_pack_bytes(doc))
