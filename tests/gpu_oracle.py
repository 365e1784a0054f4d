"""Oracle-side helpers of the GPU tests: problems rebuilt by the oracle from a configuration,
whole result packs computed by the oracle's driver, and the re-scoring of every structure a
device pack reports."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
import sqsgenerator_amd as S


def applied_species(cfg) -> np.ndarray:
    """the structure's species after the composition is distributed (structure.h:390-399)"""
    st = cfg.structure()
    applied = np.array(st.species, dtype=np.uint8)
    for sl in cfg.composition:
        it = iter(sl.sites)
        for z, n in sl.composition.items():
            for _ in range(n):
                applied[next(it)] = z
    return applied


def problem_sites(cfg, p: int) -> np.ndarray:
    if cfg.sublattice_mode == S.split:
        return np.array(cfg.composition[p].sites)
    return np.arange(len(cfg.structure()))


def oracle_inputs(cfg, job, p=0):
    """Problem p rebuilt by the oracle from the configuration; the only input taken from the
    job is the site sort order (libstdc++ std::sort of equal keys, not pinned by the
    reference)."""
    info = job.problem(p)
    st = cfg.structure()
    applied = applied_species(cfg)
    sites = problem_sites(cfg, p)
    order = info["sort_order"]
    frac = st.frac_coords[sites][order]
    z = applied[sites][order]
    dm = O.distance_matrix(st.lattice, frac)
    shells = sorted(cfg.shell_weights[p])
    sm = O.shell_matrix(dm, cfg.shell_radii[p])
    pairs = O.pairs(sm, shells)
    zs = sorted(set(z.tolist()))
    packed = np.array([zs.index(v) for v in z], dtype=np.uint8)
    sm_unsorted = O.shell_matrix(O.distance_matrix(st.lattice, st.frac_coords[sites]),
                                 cfg.shell_radii[p])
    pref = O.compute_prefactors(sm_unsorted, shells, applied[sites])
    pw = O.scaled_pair_weights(cfg.pair_weights[p], [cfg.shell_weights[p][s] for s in shells])
    return dict(pairs=pairs, packed=packed, pref=pref, pw=pw, tgt=cfg.target_objective[p],
                S=len(shells), K=len(zs), bounds=info["bounds"])


def oracle_problems(cfg, job):
    """the job's problems as the oracle's driver takes them (sorted-site order, seeds of the
    configuration); the oracle's inputs must equal the job's own"""
    probs = []
    for p in range(job.num_problems()):
        ref = oracle_inputs(cfg, job, p)
        info = job.problem(p)
        assert np.array_equal(ref["packed"], info["species_packed"])
        assert np.array_equal(ref["pairs"].astype(np.uint32), info["pairs"])
        assert np.array_equal(ref["pw"], info["pair_weights"])
        assert np.array_equal(ref["pref"], info["prefactors"])
        seed = cfg.seed[p] if p < len(cfg.seed) and cfg.seed[p] is not None else 0
        probs.append(O.Problem(ref["pairs"], ref["packed"], ref["bounds"], ref["pref"], ref["pw"],
                               ref["tgt"], seed))
    return probs


def oracle_pack(cfg, job, keep=None, max_results=None, chunk=None, iterations=None):
    """{objective: {packed species bytes}} of the oracle driver's collection: one thread
    walking the reference's blocks in order, Monte Carlo candidates by the counter-based
    definition"""
    probs = oracle_problems(cfg, job)
    _, res = O.run(probs, iterations or cfg.iterations, chunk_size=chunk or cfg.chunk_size,
                   systematic=cfg.iteration_mode == S.systematic,
                   split=cfg.sublattice_mode == S.split,
                   keep=cfg.keep if keep is None else keep,
                   max_results_per_objective=(cfg.max_results_per_objective
                                              if max_results is None else max_results),
                   threads=1, counter_stream=True)
    out = {}
    for o, sp, sub in res:
        out.setdefault(o, {})[sp.tobytes()] = sub
    return out


def device_pack(pack):
    """{objective: {packed species bytes: sub objectives}} of a device result pack"""
    out = {}
    arr = pack.arrays(species=False, ranks=False)
    for r in range(pack.num_results()):
        out.setdefault(float(arr["objective"][r]), {})[pack.species_packed(r).tobytes()] = None
    return out


def assert_packs_equal(got, ref):
    assert sorted(got) == sorted(ref), (sorted(got)[:8], sorted(ref)[:8])
    for o in ref:
        assert set(got[o]) == set(ref[o]), f"level {o}: {len(got[o])} vs {len(ref[o])}"


class Rescorer:
    """Re-scores reported structures (original site order, atomic numbers) with the oracle
    from the configuration alone: pair lists of the unsorted structure, prefactors, scaled
    pair weights; objectives bit for bit, SRO cubes exactly, ranks of interact results."""

    def __init__(self, cfg):
        self.cfg = cfg
        st = cfg.structure()
        applied = applied_species(cfg)
        self.problems = []
        for p in range(len(cfg.shell_weights)):
            sites = problem_sites(cfg, p)
            shells = sorted(cfg.shell_weights[p])
            sm = O.shell_matrix(O.distance_matrix(st.lattice, st.frac_coords[sites]),
                                cfg.shell_radii[p])
            zs = sorted(set(applied[sites].tolist()))
            self.problems.append(dict(
                sites=sites, zs=zs, pairs=O.pairs(sm, shells), S=len(shells),
                pref=cfg.prefactors[p],
                pw=O.scaled_pair_weights(cfg.pair_weights[p],
                                         [cfg.shell_weights[p][s] for s in shells]),
                tgt=cfg.target_objective[p]))

    def score(self, species_z):
        species_z = np.asarray(species_z)
        subs, sros = [], []
        for pr in self.problems:
            z = species_z[pr["sites"]]
            packed = np.array([pr["zs"].index(v) for v in z], dtype=np.uint8)
            bonds = O.count_bonds(pr["pairs"], packed, pr["S"], len(pr["zs"]))
            o, sro = O.compute_objective(bonds, pr["pref"], pr["pw"], pr["tgt"])
            subs.append(o)
            sros.append(sro)
        if len(subs) == 1:
            return subs[0], subs, sros
        tot = 0.0  # helpers::sum folds from T(0) (core/helpers/fold.h:51-53)
        for o in subs:
            tot += o
        return tot, subs, sros

    def check_pack(self, pack, limit=None):
        """every result (or the first `limit`) of a device pack; returns the count checked"""
        n = pack.num_results() if limit is None else min(limit, pack.num_results())
        if n == 0:
            return 0
        interact = self.cfg.sublattice_mode != S.split
        arr = pack.arrays(0, n, species=True, ranks=interact)
        sro = pack.sro_array(0, n)
        for r in range(n):
            o, _, sros = self.score(arr["species"][r])
            assert np.float64(o).view(np.uint64) == np.float64(arr["objective"][r]).view(
                np.uint64), (r, o, arr["objective"][r])
            assert np.array_equal(np.concatenate([s.ravel() for s in sros]), sro[r]), r
            if interact and arr["rank"][r] != np.iinfo(np.uint64).max:
                zs = self.problems[0]["zs"]
                assert int(arr["rank"][r]) == O.rank_permutation(
                    [zs.index(v) for v in arr["species"][r]]), r
        return n
