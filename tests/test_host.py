"""Host layer on the CPU: the C-ABI library loads and exports every declared entry point,
its geometry (the producer of the kernels' pair lists) equals the oracle bit for bit, the
job builder's problem inputs match an independent oracle reconstruction, and the config
parser mirrors the reference's defaults and errors (tests/test_parser.cpp)."""
import ctypes as C
import json
import re
from pathlib import Path

import numpy as np
import pytest

import workloads as W
from oracle import oracle as O
import sqsgenerator_amd as S
import sqsgenerator_amd._lib as L
from sqsgenerator_amd.core import Job, ParseError, parse_config

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = "".join(p.read_text() for p in (ROOT / "include").glob("*.h"))
    return sorted(set(re.findall(r"SQSG_API\s+[\w\s\*]+?\b(sqsg_\w+)\s*\(", text)))


def test_abi_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) <= set(L.SIGNATURES), sorted(set(syms) - set(L.SIGNATURES))


def test_error_channel():
    lib = L.lib()
    assert lib.sqsg_distance_matrix(None, None, 0, None) == L.ERR_INVALID_ARGUMENT
    assert b"must not be NULL" in lib.sqsg_last_error()
    c = np.zeros(16, dtype=np.uint8)
    c[8:] = 1
    assert lib.sqsg_unrank_permutation(L.ptr(c, L.u8p), 16, b"12871") == L.ERR_OUT_OF_RANGE


STRUCTS = {
    "case1": lambda: (lambda d: (np.array(d["lattice"]), np.array(d["frac_coords"])))(
        json.loads((ROOT / "tests/golden/case1.structure.json").read_text())),
    "fcc3": lambda: O.supercell(np.eye(3) * 4.05, W.FCC_BASIS, [1] * 4, 3, 3, 3)[:2],
    "bcc4": lambda: O.supercell(np.eye(3) * 3.165, W.BCC_BASIS, [1] * 2, 4, 4, 4)[:2],
    "hcp4": lambda: O.supercell(
        [[2.95, 0, 0], [-1.475, 2.95 * np.sqrt(3) / 2, 0], [0, 0, 1.633 * 2.95]],
        [[1 / 3, 2 / 3, .25], [2 / 3, 1 / 3, .75]], [1] * 2, 4, 4, 4)[:2],
}


@pytest.mark.parametrize("name", list(STRUCTS))
def test_geometry_matches_oracle(name):
    lat, frac = STRUCTS[name]()
    dm = S.distance_matrix(lat, frac)
    assert np.array_equal(dm, O.distance_matrix(lat, frac))
    for args in ((np.finfo(float).eps, 1e-9), (1e-3, 1e-5)):
        assert np.array_equal(S.distances_naive(dm, *args), O.distances_naive(dm, *args))
    try:
        peak = O.distances_histogram(dm)
    except ValueError:
        with pytest.raises(IndexError):
            S.distances_histogram(dm)
        peak = None
    if peak is not None:
        assert np.array_equal(S.distances_histogram(dm), peak)
    radii = O.distances_naive(dm)[:8]
    sm = S.shell_matrix(dm, radii)
    assert np.array_equal(sm, O.shell_matrix(dm, radii))
    z = np.arange(len(frac)) % 3 + 20
    shells = [1, 2, 3]
    assert np.array_equal(S.prefactors(sm, shells, z), O.compute_prefactors(sm, shells, z))


def test_supercell_matches_oracle():
    st = S.Structure(np.diag([4.05] * 3), W.FCC_BASIS, ["Al"] * 4).supercell(3, 2, 4)
    Lo, fo, zo = O.supercell(np.diag([4.05] * 3), W.FCC_BASIS, [13] * 4, 3, 2, 4)
    assert np.array_equal(st.lattice, Lo) and np.array_equal(st.frac_coords, fo)
    assert st.species == zo.tolist()


def test_permutations_match_oracle():
    rng = np.random.default_rng(3)
    for _ in range(50):
        conf = rng.permutation(np.repeat(np.arange(4, dtype=np.uint8), rng.integers(1, 6, 4)))
        assert S.num_permutations(conf) == O.num_permutations(conf)
        r = S.rank_permutation(conf)
        assert r == O.rank_permutation(conf)
        assert np.array_equal(S.unrank_permutation(np.sort(conf), r), conf)
    big = np.repeat(np.arange(5, dtype=np.uint8), 52)  # multinomial far beyond 128 bits
    assert S.num_permutations(big) > 2 ** 400
    assert S.rank_permutation(big) == 1 and S.rank_permutation(big[::-1]) == S.num_permutations(big)


CONFIGS = {
    "fcc32": W.plain_mc_32, "fcc108": W.ternary_108, "bcc128": W.quaternary_bcc_128,
    "hcp128": W.hcp_128, "rocksalt216": W.rocksalt_split_216, "fcc256": W.quinary_256,
    "test16_partial": lambda: dict(W.reference_test_16(), structure=dict(
        W.reference_test_16()["structure"], species=[1, 2] * 2, supercell=[2, 2, 2])),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_job_inputs_match_oracle(name):
    """the job's pair lists, prefactors and scaled weights (optimization_config_data) equal an
    independent oracle reconstruction; only the site sort order is taken from the job"""
    cfg = parse_config(CONFIGS[name]())
    job = Job(cfg)
    st = cfg.structure()
    applied = np.array(st.species, dtype=np.uint8)
    for sl in cfg.composition:
        it = iter(sl.sites)
        for z, n in sl.composition.items():
            for _ in range(n):
                applied[next(it)] = z
    for p in range(job.num_problems()):
        info = job.problem(p)
        sites = (np.array(cfg.composition[p].sites) if cfg.sublattice_mode == S.split
                 else np.arange(len(applied)))
        order = info["sort_order"]
        assert sorted(order.tolist()) == list(range(len(sites)))
        shells = sorted(cfg.shell_weights[p])
        sm = O.shell_matrix(O.distance_matrix(st.lattice, st.frac_coords[sites][order]),
                            cfg.shell_radii[p])
        assert np.array_equal(info["pairs"].astype(np.uint64), O.pairs(sm, shells))
        sm0 = O.shell_matrix(O.distance_matrix(st.lattice, st.frac_coords[sites]),
                             cfg.shell_radii[p])
        assert np.array_equal(info["prefactors"],
                              O.compute_prefactors(sm0, shells, applied[sites]))
        assert np.array_equal(info["pair_weights"], O.scaled_pair_weights(
            cfg.pair_weights[p], [cfg.shell_weights[p][s] for s in shells]))
        z = applied[sites][order]
        assert [info["species_z"][v] for v in info["species_packed"]] == z.tolist()
        # shuffle windows: one per interact sublattice, in sublattice order
        for (lo, hi), sl in zip(info["bounds"], cfg.composition if p == 0 and
                                cfg.sublattice_mode == S.interact else [cfg.composition[p]]):
            if cfg.sublattice_mode == S.interact:
                assert sorted(order[lo:hi].tolist()) == sl.sites
    job.close()


def test_parse_defaults():
    d = W.reference_test_16(iteration_mode="random")
    del d["pair_weights"]
    cfg = parse_config(d)
    assert cfg.iterations == 500000 and cfg.chunk_size == 100000 and cfg.keep == 1
    # default shell weights 1/i for every detected shell (test_parser.cpp:337-359)
    assert cfg.shell_weights[0] == {i: 1.0 / i for i in range(1, len(cfg.shell_radii[0]))}
    # default pair weights J - I (arrays.h:103-107), target 0 (arrays.h:180-190)
    assert np.array_equal(cfg.pair_weights[0][0], 1 - np.eye(2))
    assert not cfg.target_objective[0].any()
    assert cfg.seed == [None]


@pytest.mark.parametrize("patch,key", [
    ({"shell_weights": {"0": 1.0}}, "shell_weights"),
    ({"shell_weights": {"200": 1.0}}, "shell_weights"),
    ({"iteration_mode": "systematic", "seed": 4}, "seed"),
    ({"iteration_mode": "random", "seed": 42, "thread_config": 4}, "seed"),
    ({"iteration_mode": "systematic", "sublattice_mode": "split"}, "iteration_mode"),
    ({"composition": {"sites": "H", "Ti": 8, "Al": 7}}, "sites"),
    ({"composition": {"sites": "Xx", "Ti": 8}}, "sites"),
    ({"iteration_mode": "random", "iterations": 10, "chunk_size": 20}, "chunk_size"),
    ({"keep": 0}, "keep"),
    ({"bogus": 1}, "bogus"),
    ({"pair_weights": [[1, 2, 3]]}, "pair_weights"),
])
def test_parse_errors(patch, key):
    d = W.reference_test_16()
    d.update(patch)
    err = parse_config(d)
    assert isinstance(err, ParseError), err
    assert err.key == key


def test_seed_single_thread_reproducible_config():
    """python/tests/test_core.py:291-301"""
    d = W.reference_test_16(iteration_mode="random", iterations=500, seed=42)
    a, b = parse_config(d), parse_config(d)
    assert a.seed == b.seed == [42]


def test_kernel_shape_follows_the_applicability_rules():
    """the launch shape a run would use, resolved on the host (no device needed): AUTO takes
    the dense form for one problem of at most 1024 sites; the paired form takes Monte Carlo
    jobs of at most 256 sites; the lane kernel is sized by K"""
    from sqsgenerator_amd.core import Job
    job = Job(parse_config(W.quinary_256(iterations=1000)))
    assert job.kernel_shape() == "dense<8,2,2,3> columns 4"
    job.set_kernel(4)
    assert job.kernel_shape() == "paired<8,2,9>"
    job.set_kernel(2)
    assert job.kernel_shape() == "lane<5>"
    job.set_kernel(1)
    assert job.kernel_shape() == "wave"
    job.close()
    big = Job(parse_config(W.quinary_864(iterations=1000)))
    assert big.kernel_shape() == "dense<27,1,2,9> columns 4"
    big.set_kernel(4)
    with pytest.raises(ValueError):
        big.kernel_shape()
    big.close()


@pytest.mark.parametrize("make,K,shape", [
    (lambda: W.quinary_256(iterations=1000), 5, "dense<8,2,2,3> columns 4"),
    (lambda: W.ternary_108(iterations=1000), 3, "dense<4,2,2,5> columns 2"),
    (lambda: W.rocksalt_split_216(iterations=1000), 2, "dense<4,2,2,5> columns 1+1"),
    # B and X windows with 6 and 2 neighbours at a/2, the A sites (Sr) fixed
    (lambda: W.perovskite_bx(), 5, "dense<6,2,2,3> columns 4"),
    # one window over sites with different neighbour counts: every species keeps a column
    (lambda: W.perovskite_bx(mixed=True), 4, "dense<6,2,2,3> columns 4"),
])
def test_dense_columns_per_candidate(make, K, shape):
    """the dense search kernels leave the last species' one-hot column out wherever every
    shuffle window's sites have equal neighbour counts in each shell"""
    from sqsgenerator_amd.core import Job
    job = Job(parse_config(make()))
    assert len(set(job.problem(0)["species_packed"].tolist())) == K
    assert job.kernel_shape() == shape
    job.close()


@pytest.mark.parametrize("sites,cell,bound", [(32, (2, 2, 2), 1), (96, (2, 3, 4), 4),
                                              (192, (4, 4, 3), 6), (256, (4, 4, 4), 8),
                                              (384, (4, 4, 6), 12), (432, (3, 4, 9), 14),
                                              (500, (5, 5, 5), 16), (600, (5, 5, 6), 20),
                                              (768, (4, 8, 6), 24), (864, (6, 6, 6), 27),
                                              (1024, (8, 8, 4), 32)])
def test_dense_tile_bound_by_cell_size(sites, cell, bound):
    """the dense kernel runs the smallest tile bound that holds the cell's 32-site tiles"""
    from sqsgenerator_amd.core import Job
    comp = {"Fe": sites // 2, "Co": sites - sites // 2}
    job = Job(parse_config(W.fcc(4.05, cell, comp, 3, iterations=1000, thread_config=1)))
    assert job.kernel_shape().startswith(f"dense<{bound},")
    job.close()
    sys16 = W.reference_test_16()
    del sys16["pair_weights"]
    syst = Job(parse_config(sys16))
    syst.set_kernel(4)
    with pytest.raises(ValueError):
        syst.kernel_shape()
    syst.close()


def test_pruning_needs_non_negative_pair_weights():
    from sqsgenerator_amd.core import Job
    job = Job(parse_config(W.quinary_256(iterations=1000)))
    job.set_pruning(True)
    job.set_pruning(False)
    job.close()
    d = W.reference_test_16()
    d["pair_weights"] = [[0.0, -1.0], [-1.0, 0.0]]
    cfg = parse_config(d)
    assert not isinstance(cfg, ParseError), cfg
    job = Job(cfg)
    with pytest.raises(ValueError):
        job.set_pruning(True)
    job.set_pruning(False)
    job.close()
