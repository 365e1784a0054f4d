"""configuration<float> ("prec": "single", io/config/combined.h:283-295, sqs.h:591-634).

CPU: the oracle's float restatements against closed forms, invariants and the double
functions, and the product's float geometry against the oracle.  GPU: explicit scoring and
whole runs of float jobs re-scored by the oracle's float functions, bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import workloads as W
from gpu_oracle import applied_species, problem_sites
from oracle import oracle as O
from sqsgenerator_amd import _lib as L
from sqsgenerator_amd import core

F32_EPS = float(np.finfo(np.float32).eps)


def single(doc: dict) -> dict:
    d = dict(doc)
    d["prec"] = "single"
    return d


def f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float32)


# ---- oracle, float restatements ---------------------------------------------------------

def test_oracle_objective_f32_closed_forms():
    rng = np.random.default_rng(3)
    S_, K = 3, 4
    bonds = rng.integers(0, 200, size=(S_, K, K)).astype(np.uint64)
    zero = np.zeros((S_, K, K))
    # zero pair weights: every term vanishes (test_core.py:39,59)
    assert O.compute_objective_f32(bonds, rng.random((S_, K, K)), zero, zero)[0] == 0.0
    # unit prefactors, weights and a zero target: sum of |1 - pair bonds|, exact in float
    ones = np.ones((S_, K, K))
    o, sro = O.compute_objective_f32(bonds, ones, ones, zero)
    want = 0
    for s in range(S_):
        for a in range(K):
            want += abs(1 - int(bonds[s, a, a]))
            for b in range(a + 1, K):
                pb = int(bonds[s, a, b] + bonds[s, b, a])
                want += abs(1 - pb)
                assert sro[s, a, b] == sro[s, b, a] == np.float32(1 - pb)
    assert o == np.float32(want)


def test_oracle_objective_f32_agrees_with_double():
    """the float objective is the double one up to float rounding of each operation"""
    rng = np.random.default_rng(4)
    for _ in range(200):
        S_, K = rng.integers(1, 8), rng.integers(2, 7)
        bonds = rng.integers(0, 500, size=(S_, K, K)).astype(np.uint64)
        pref = f32(rng.random((S_, K, K)) * 0.01)
        pw = f32(rng.random((S_, K, K)))
        tgt = f32(rng.random((S_, K, K)) - 0.5)
        of, sf = O.compute_objective_f32(bonds, pref, pw, tgt)
        od, sd = O.compute_objective(bonds, pref.astype(np.float64), pw.astype(np.float64),
                                     tgt.astype(np.float64))
        terms = S_ * K * (K + 1) // 2
        assert abs(float(of) - od) <= 4 * terms * F32_EPS * (abs(od) + 1.0)
        assert np.allclose(sf, sd, rtol=0, atol=8 * F32_EPS * (np.abs(sd).max() + 1))


def test_oracle_geometry_f32_fcc_coordination():
    """FCC 3x3x3: shells of 12, 6, 24, 12, 12, 8 neighbours from the float shell matrix
    (test_structure.cpp:219-237); prefactors 1 / (M x_a x_b N) in float"""
    a = 4.05
    lat = np.diag([a, a, a])
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    cells = np.array([[x, y, z] for x in range(3) for y in range(3) for z in range(3)])
    frac = ((cells[:, None, :] + basis[None]) / 3.0).reshape(-1, 3)
    L = np.diag([3 * a] * 3)
    dm = O.distance_matrix_f32(L, frac)
    dmd = O.distance_matrix(L, f32(frac).astype(np.float64))
    assert np.allclose(dm, dmd, rtol=4 * F32_EPS, atol=0)
    radii = f32([0.0] + [a * np.sqrt(k / 2.0) for k in (1, 2, 3, 4, 5, 6)])
    sm = O.shell_matrix_f32(dm, radii, atol=1e-3, rtol=1e-5)
    counts = [int((sm[0] == s).sum()) for s in range(1, 7)]
    assert counts == [12, 6, 24, 12, 12, 8]
    z = np.array([13] * 54 + [22] * 54, dtype=np.uint8)
    pref = O.compute_prefactors_f32(sm, [1, 2, 3], z)
    for s, M in enumerate((12, 6, 24)):
        want = np.float32(1.0) / (np.float32(M) * np.float32(0.5) * np.float32(0.5) *
                                  np.float32(108))
        assert np.all(pref[s] == want)
    w = O.scaled_pair_weights_f32(np.ones((3, 2, 2)), [1.0, 0.5, 1.0 / 3.0])
    assert w[2, 0, 1] == np.float32(1.0) * np.float32(1.0 / 3.0)


# ---- product host geometry in float ------------------------------------------------------

@pytest.mark.parametrize("make", [W.quinary_256, W.hcp_128, W.rocksalt_split_216])
def test_float_config_geometry_matches_oracle(make):
    cfg = core.parse_config(single(make(iterations=1000)))
    assert isinstance(cfg, core.SqsConfiguration) and cfg.prec == "single"
    st = cfg.structure()
    assert st.single and np.array_equal(st.frac_coords, f32(st.frac_coords))
    applied = applied_species(cfg)
    for p in range(len(cfg.shell_weights)):
        sites = problem_sites(cfg, p)
        dm = O.distance_matrix_f32(st.lattice, st.frac_coords[sites])
        assert np.array_equal(core.distance_matrix(st.lattice, st.frac_coords[sites], True), dm)
        sm = O.shell_matrix_f32(dm, cfg.shell_radii[p])
        assert np.array_equal(core.shell_matrix(dm, cfg.shell_radii[p], core.EPS_F32, 1e-9,
                                                True), sm)
        shells = sorted(cfg.shell_weights[p])
        assert all(np.float32(w) == w for w in cfg.shell_weights[p].values())
        assert np.array_equal(cfg.prefactors[p],
                              O.compute_prefactors_f32(sm, shells, applied[sites]))
        assert np.array_equal(cfg.target_objective[p], f32(cfg.target_objective[p]))


def test_float_radii_and_weights_are_float():
    d = W.quinary_256(iterations=1000)
    cd, cf = core.parse_config(d), core.parse_config(single(d))
    rf = np.asarray(cf.shell_radii[0])
    assert np.array_equal(rf, f32(rf)) and not np.array_equal(rf, cd.shell_radii[0])
    assert cf.shell_weights[0][3] == float(np.float32(1.0) / np.float32(3.0))
    assert cd.shell_weights[0][3] == 1.0 / 3.0


# ---- device -------------------------------------------------------------------------------

class FloatRescorer:
    """re-scores reported structures with the oracle's float functions from the
    configuration alone"""

    def __init__(self, cfg):
        self.cfg = cfg
        st = cfg.structure()
        applied = applied_species(cfg)
        self.problems = []
        for p in range(len(cfg.shell_weights)):
            sites = problem_sites(cfg, p)
            shells = sorted(cfg.shell_weights[p])
            sm = O.shell_matrix_f32(O.distance_matrix_f32(st.lattice, st.frac_coords[sites]),
                                    cfg.shell_radii[p])
            zs = sorted(set(applied[sites].tolist()))
            self.problems.append(dict(
                sites=sites, zs=zs, pairs=O.pairs(sm, shells), S=len(shells),
                pref=O.compute_prefactors_f32(sm, shells, applied[sites]),
                pw=O.scaled_pair_weights_f32(cfg.pair_weights[p],
                                             [cfg.shell_weights[p][s] for s in shells]),
                tgt=f32(cfg.target_objective[p])))

    def score(self, species_z):
        subs, sros = [], []
        for pr in self.problems:
            z = np.asarray(species_z)[pr["sites"]]
            packed = np.array([pr["zs"].index(v) for v in z], dtype=np.uint8)
            bonds = O.count_bonds(pr["pairs"], packed, pr["S"], len(pr["zs"]))
            o, sro = O.compute_objective_f32(bonds, pr["pref"], pr["pw"], pr["tgt"])
            subs.append(o)
            sros.append(sro)
        if len(subs) == 1:
            return subs[0], sros
        tot = np.float32(0.0)  # helpers::sum folds from T(0), in float
        for o in subs:
            tot = np.float32(tot + o)
        return tot, sros

    def check_pack(self, pack):
        n = pack.num_results()
        assert n > 0
        arr = pack.arrays(0, n, species=True, ranks=False)
        sro = pack.sro_array(0, n)
        for r in range(n):
            o, sros = self.score(arr["species"][r])
            assert np.float64(o) == arr["objective"][r], (r, o, arr["objective"][r])
            assert np.array_equal(np.concatenate([s.ravel() for s in sros]).astype(np.float64),
                                  sro[r]), r
        return n


@pytest.mark.gpu
@pytest.mark.parametrize("make", [W.quinary_256, W.hcp_128])
def test_float_explicit_scoring_bit_exact(gpu, make):
    cfg = core.parse_config(single(make(iterations=4096)))
    job = core.Job(cfg)
    info = job.problem(0)
    resc = FloatRescorer(cfg)
    pr = resc.problems[0]
    assert np.array_equal(info["pair_weights"], pr["pw"].astype(np.float64))
    rng = np.random.default_rng(5)
    cands = np.stack([rng.permutation(info["species_packed"]) for _ in range(256)])
    for kernel in (L.KERNEL_SPARSE_WAVE, L.KERNEL_SPARSE_LANE):
        bonds, obj, sro = job.score(cands, 0, kernel=kernel)
        for b in range(len(cands)):
            ref = O.count_bonds(info["pairs"].astype(np.uint64), cands[b], info["num_shells"],
                                info["num_species"])
            assert np.array_equal(bonds[b].astype(np.uint64), ref)
            o, s = O.compute_objective_f32(ref, info["prefactors"], info["pair_weights"],
                                           info["target"])
            assert obj[b] == np.float64(o), (kernel, b)
            assert np.array_equal(sro[b], s.astype(np.float64)), (kernel, b)
    with pytest.raises(Exception):
        job.score(cands, 0, kernel=L.KERNEL_DENSE_MFMA)
    job.close()


@pytest.mark.gpu
@pytest.mark.parametrize("make,iterations", [(W.quinary_256, 200_000), (W.ternary_108, 500_000),
                                             (W.rocksalt_split_216, 200_000)])
def test_float_runs_rescored(gpu, make, iterations):
    cfg = core.parse_config(single(make(iterations=iterations)))
    pack = core.optimize(cfg)
    assert FloatRescorer(cfg).check_pack(pack) > 0
    # the same configuration in double: same bonds, different arithmetic
    dpack = core.optimize(core.parse_config(make(iterations=iterations)))
    assert dpack.num_results() > 0


@pytest.mark.gpu
def test_float_job_batched_with_double_job(gpu):
    cf = core.parse_config(single(W.ternary_108(iterations=300_000)))
    cd = core.parse_config(W.quinary_256(iterations=300_000))
    alone = core.optimize(cf)
    jobs = [core.Job(cf), core.Job(cd)]
    core.batch_run(jobs)
    got = jobs[0].results()
    a = alone.arrays(0, alone.num_results(), species=True, ranks=False)
    g = got.arrays(0, got.num_results(), species=True, ranks=False)
    assert np.array_equal(a["objective"], g["objective"])
    assert np.array_equal(a["species"], g["species"])
    FloatRescorer(cf).check_pack(got)
    for j in jobs:
        j.close()


@pytest.mark.gpu
def test_float_pack_round_trip(gpu):
    """a float run's pack through the wire form and load_result_pack(..., "single")"""
    cfg = core.parse_config(single(W.ternary_108(iterations=100_000)))
    job = core.Job(cfg)
    job.run()
    pack = job.results()
    loaded = core.load_result_pack(pack.bytes(), "single")
    assert loaded.config.prec == "single"
    la, lb = list(pack), list(loaded)
    assert [o for o, _ in la] == [o for o, _ in lb]
    assert all(np.float64(np.float32(o)) == o for o, _ in lb)
    for (_, ra), (_, rb) in zip(la, lb):
        sa = sorted((tuple(r.species), tuple(r.sro().ravel())) for r in ra)
        sb = sorted((tuple(r.species), tuple(r.sro().ravel())) for r in rb)
        assert sa == sb
    job.close()
