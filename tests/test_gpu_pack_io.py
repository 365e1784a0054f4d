"""Result packs of real runs through the reference's msgpack wire format (bytes ->
load_result_pack), after python/tests/test_core.py:130-177 test_structure_pack_io."""
import numpy as np
import pytest

import workloads as W
import sqsgenerator_amd as S
from sqsgenerator_amd.core import Job, parse_config

pytestmark = pytest.mark.gpu


def _by_species(pack):
    return {"".join(map(str, r.species)): r for _, rs in pack for r in rs}


def test_systematic_pack_round_trip(gpu, tmp_path):
    """the 16-site Ti8Al8 systematic run keeps all 12870 configurations (pair weights 0):
    every one survives the wire format with its objective, SRO, rank and structure"""
    d = W.reference_test_16()
    print("[pack-io] systematic 16", flush=True)
    results = S.optimize(d)
    data = results.bytes()
    path = tmp_path / "pack.mpack"
    path.write_bytes(data)
    for loaded in (S.load_result_pack(data), S.load_result_pack(str(path))):
        ra, rb = _by_species(results), _by_species(loaded)
        assert len(ra) == len(rb) == loaded.num_results() == results.num_results() == 12870
        assert not (set(ra) ^ set(rb))
        for k, a in ra.items():
            b = rb[k]
            assert a.objective == b.objective
            assert np.array_equal(a.sro(), b.sro())
            assert a.rank() == b.rank()
            assert np.array_equal(a.structure().frac_coords, b.structure().frac_coords)


def test_mc_pack_round_trip(gpu):
    d = W.plain_mc_32(iterations=200_000)
    d["chunk_size"] = 50_000
    print("[pack-io] mc 32", flush=True)
    job = Job(parse_config(d))
    job.run()
    pack = job.results()
    loaded = S.load_result_pack(pack.bytes())
    la, lb = list(pack), list(loaded)
    assert [o for o, _ in la] == [o for o, _ in lb]
    for (_, ra), (_, rb) in zip(la, lb):
        sa = sorted((tuple(r.species), r.rank(), tuple(r.sro().ravel())) for r in ra)
        sb = sorted((tuple(r.species), r.rank(), tuple(r.sro().ravel())) for r in rb)
        assert sa == sb
    assert loaded.statistics["finished"] == pack.statistics["finished"]
    job.close()


def test_split_pack_round_trip(gpu):
    d = W.rocksalt_split_216(iterations=50_000)
    d["chunk_size"] = 50_000
    print("[pack-io] split 216", flush=True)
    job = Job(parse_config(d))
    job.run()
    pack = job.results()
    loaded = S.load_result_pack(pack.bytes())
    la, lb = list(pack), list(loaded)
    assert [o for o, _ in la] == [o for o, _ in lb]
    for (_, ra), (_, rb) in zip(la, lb):
        key = lambda r: tuple(r.species)
        for a, b in zip(sorted(ra, key=key), sorted(rb, key=key)):
            assert a.species == b.species
            for x, y in zip(a.sublattices, b.sublattices):
                assert x.objective == y.objective
                assert np.array_equal(x.sro(), y.sro())
                assert x.rank() == y.rank()
    job.close()
