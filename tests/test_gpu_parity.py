"""GPU parity: the HIP path (through the C ABI) against the oracle on the same inputs.

Integer work must be bit-exact (bond counts, ranks); objectives are compared bit-for-bit
(both sides evaluate the reference's term order without contraction), which is stricter
than a 1e-12 relative tolerance.
"""
import numpy as np
import pytest

import workloads as W
from oracle import oracle as O
import sqsgenerator_amd as S
from gpu_oracle import Rescorer, assert_packs_equal, device_pack, oracle_inputs, oracle_pack
from sqsgenerator_amd.core import Job, parse_config

pytestmark = pytest.mark.gpu
REL_TOL = 1e-12


def random_candidates(packed, bounds, B, rng):
    out = np.repeat(packed[None, :], B, axis=0)
    for b in range(B):
        for lo, hi in bounds:
            out[b, lo:hi] = rng.permutation(out[b, lo:hi])
    return out


WORKLOADS = {
    "fcc32_3shells": lambda: W.plain_mc_32(),
    "fcc108_ternary": lambda: W.ternary_108(),
    "bcc128_quaternary": lambda: W.quaternary_bcc_128(),
    "hcp128_binary": lambda: W.hcp_128(),
    "rocksalt216_split": lambda: W.rocksalt_split_216(),
    "fcc256_quinary": lambda: W.quinary_256(),
    "fcc864_quinary": lambda: W.quinary_864(),
    "fcc64_ternary": lambda: W.ternary_64(),
    "fcc500_quaternary": lambda: W.quaternary_500(),
}


KERNELS = {"wave": 1, "lane": 2, "dense": 3}
SINGLE_PROBLEM = [n for n in WORKLOADS if not n.endswith("_split")]
DENSE = list(WORKLOADS)  # dense kernel domain: split jobs too (sublattices up to 256 sites)
PAIRED = ["fcc32_3shells", "fcc108_ternary", "bcc128_quaternary", "hcp128_binary",
          "fcc256_quinary"]  # paired dense form: single problem, at most 256 sites


@pytest.mark.parametrize("kernel", list(KERNELS))
@pytest.mark.parametrize("name", list(WORKLOADS))
def test_score_batch_bit_exact(gpu, name, kernel):
    print(f"[score_batch] {name} {kernel}", flush=True)
    cfg = parse_config(WORKLOADS[name]())
    job = Job(cfg)
    rng = np.random.default_rng(20261203)
    for p in range(job.num_problems()):
        ref = oracle_inputs(cfg, job, p)
        info = job.problem(p)
        assert np.array_equal(info["pairs"].astype(np.uint64), ref["pairs"]), "pair list"
        assert np.array_equal(info["species_packed"], ref["packed"])
        assert np.array_equal(info["prefactors"], ref["pref"])
        assert np.array_equal(info["pair_weights"], ref["pw"])
        B = 64 if info["num_sites"] < 500 else 16
        cands = random_candidates(ref["packed"], ref["bounds"], B, rng)
        cands[0] = ref["packed"]  # the sorted configuration itself
        bonds, obj, sro = job.score(cands, p, kernel=KERNELS[kernel])
        for b in range(B):
            rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
            assert np.array_equal(bonds[b].astype(np.uint64), rb), f"bonds of candidate {b}"
            ro, rsro = O.compute_objective(rb, ref["pref"], ref["pw"], ref["tgt"])
            assert obj[b] == ro or abs(obj[b] - ro) <= REL_TOL * abs(ro), (b, obj[b], ro)
            assert obj[b].view(np.uint64) == np.float64(ro).view(np.uint64), "bit-exact objective"
            assert np.array_equal(sro[b], rsro)
            # bond totals per shell equal the shell's pair count (python/tests/test_core.py:245-248)
            assert np.array_equal(bonds[b].sum(axis=(1, 2)),
                                  np.bincount(ref["pairs"][:, 2].astype(int), minlength=ref["S"]))
    job.close()


@pytest.mark.parametrize("name", SINGLE_PROBLEM)
def test_score_dense_objective_only(gpu, name):
    """explicit scoring of objectives and SRO cubes without the ordered bond counts
    (a null bonds buffer): bit-exact like the full output"""
    print(f"[score_dense_objective_only] {name}", flush=True)
    cfg = parse_config(WORKLOADS[name]())
    job = Job(cfg)
    ref = oracle_inputs(cfg, job, 0)
    rng = np.random.default_rng(20261204)
    B = 96 if job.problem(0)["num_sites"] < 500 else 16
    cands = random_candidates(ref["packed"], ref["bounds"], B, rng)
    cands[1] = ref["packed"]
    bonds, obj, sro = job.score(cands, 0, kernel=KERNELS["dense"], bonds=False)
    assert bonds is None
    for b in range(B):
        rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
        ro, rsro = O.compute_objective(rb, ref["pref"], ref["pw"], ref["tgt"])
        assert obj[b].view(np.uint64) == np.float64(ro).view(np.uint64), (b, obj[b], ro)
        assert np.array_equal(sro[b], rsro)
    job.close()


def test_systematic_all_results(gpu):
    """python/tests/test_core.py:45-67: pair_weights 0 -> every permutation is optimal"""
    cfg = parse_config(W.reference_test_16())
    assert cfg.iterations == 12870
    res = S.optimize(cfg)
    assert res.num_results() == 12870
    ranks = set(range(1, 12871))
    for _, rset in res:
        for sol in rset:
            assert sol.objective == 0.0
            assert sorted(sol.species) == [13] * 8 + [22] * 8
            ranks.discard(int(sol.rank()))
    assert not ranks


def test_systematic_partial_sublattice(gpu):
    """python/tests/test_core.py:82-128: only the H sites are permuted, He sites stay"""
    d = W.reference_test_16()
    d["structure"]["species"] = [1, 2] * 2
    d["structure"]["supercell"] = [2, 2, 2]
    cfg = parse_config(d)
    res = S.optimize(cfg)
    assert res.num_results() == cfg.iterations == 12870
    st = cfg.structure()
    he = {i for i, z in enumerate(st.species) if z == 2}
    for _, rset in res:
        for sol in rset:
            assert {i for i, z in enumerate(sol.species) if z == 2} == he
            assert sorted(z for z in sol.species if z != 2) == [13] * 8 + [22] * 8


def test_mc_plain_32_rescored(gpu):
    """configuration (1): 1e6 MC candidates; the pack equals the oracle driver's over the
    same candidate stream, and every reported structure is re-scored from the configuration
    (objective bits, SRO, rank)"""
    cfg = parse_config(W.plain_mc_32(iterations=1_000_000))
    job = Job(cfg)
    job.run()
    res = job.results()
    assert res.statistics["finished"] == 1_000_000
    assert Rescorer(cfg).check_pack(res) >= 1
    assert_packs_equal(device_pack(res), oracle_pack(cfg, job))
    job.close()


def test_mc_tie_set_complete_16(gpu):
    """MC over the 12870-candidate space finds every configuration of the best two levels
    (1e6 draws: miss probability per configuration e^-77); every candidate at or below the
    final threshold is inserted, so the pack's two best levels are complete"""
    base = W.reference_test_16()
    del base["pair_weights"]
    sys_cfg = parse_config(base)
    mc = dict(base, iteration_mode="random", iterations=1_000_000, chunk_size=100_000)
    mc_cfg = parse_config(mc)
    # oracle: exhaustive over the same problem (identity site order)
    st = sys_cfg.structure()
    shells = sorted(sys_cfg.shell_weights[0])
    sm = O.shell_matrix(O.distance_matrix(st.lattice, st.frac_coords), sys_cfg.shell_radii[0])
    pairs = O.pairs(sm, shells)
    z0 = np.array([13] * 8 + [22] * 8, dtype=np.uint8)
    packed0 = (z0 == 22).astype(np.uint8)
    pw = O.scaled_pair_weights(sys_cfg.pair_weights[0],
                               [sys_cfg.shell_weights[0][s] for s in shells])
    prob = O.Problem(pairs, packed0, [0, 16], sys_cfg.prefactors[0], pw,
                     sys_cfg.target_objective[0])
    _, ref = O.run([prob], 12870, systematic=True, keep=1)
    ref_levels = {}
    for o, sp, _ in ref:
        ref_levels.setdefault(o, set()).add(tuple(np.where(sp == 1, 22, 13).tolist()))
    best = sorted(ref_levels)[:mc_cfg.keep + 1]
    res = S.optimize(mc_cfg)
    got = {o: {tuple(s.species) for s in rs} for o, rs in res}
    assert sorted(got)[:mc_cfg.keep + 1] == best
    for o in best:
        assert got[o] == ref_levels[o], f"level {o}"


def _pack_signature(pack):
    return [(o, sorted((tuple(r.species), r.candidate) for r in rs)) for o, rs in pack]


@pytest.mark.parametrize("name", ["fcc32_3shells", "fcc256_quinary", "rocksalt216_split",
                                  "hcp128_binary", "fcc864_quinary"])
def test_variants_agree_mc(gpu, name):
    """candidate t is the same configuration in every kernel (counter-based streams), so the
    wave, lane and dense kernels must return identical result packs for the same seed"""
    print(f"[variants] {name}", flush=True)
    d = WORKLOADS[name]()
    d["iterations"] = 200_000 if name != "fcc864_quinary" else 20_000
    d["chunk_size"] = min(d["iterations"], 50_000)
    packs = []
    kernels = (1, 2, 3, 4) if name in PAIRED else (1, 2, 3)
    for kernel in kernels:
        job = Job(parse_config(d))
        job.set_kernel(kernel)
        job.run()
        packs.append(_pack_signature(job.results()))
        assert job.timing()[2] == d["iterations"]
        job.close()
    for other in packs[1:]:
        assert other == packs[0]
    job = Job(parse_config(d))
    job.run()
    assert Rescorer(job.config).check_pack(job.results()) > 0
    job.close()


def test_variants_agree_systematic(gpu):
    d = W.reference_test_16()
    del d["pair_weights"]
    packs = []
    for kernel in (1, 2, 3):
        job = Job(parse_config(d))
        job.set_kernel(kernel)
        job.set_launch_size(1000)
        job.run()
        packs.append(_pack_signature(job.results()))
        job.close()
    assert packs[0] == packs[1] == packs[2]


def test_systematic_32_optimal_rank_set(gpu):
    """all C(32,16) = 601,080,390 permutations of the 32-site A16B16 cell (2 shells): same
    optimal objective and the same set of optimal reported ranks as the oracle's exhaustive
    search (tests/golden/systematic_fcc32.json, 17,271,296 optimal ranks)"""
    import hashlib
    import json
    from pathlib import Path

    golden = json.loads((Path(__file__).parent / "golden" / "systematic_fcc32.json").read_text())
    cfg = parse_config(W.systematic_32())
    assert cfg.iterations == golden["total"]
    assert np.array_equal(cfg.shell_radii[0], golden["radii"])
    job = Job(cfg)
    job.run()
    assert job.timing()[2] == golden["total"]
    pack = job.results()
    assert float(pack.level_objectives[0]).hex() == golden["best_objective_hex"]
    assert float(pack.level_objectives[1]).hex() == golden["second_objective_hex"]
    n_best = int(pack.level_start[1])
    assert n_best == golden["num_best"]
    assert int(pack.level_start[2] - pack.level_start[1]) == golden["num_second"]
    ranks = np.sort(pack.arrays(0, n_best, species=False, ranks=True)["rank"])
    assert ranks[:64].tolist() == golden["best_ranks_head"]
    assert hashlib.sha256(ranks.astype("<u8").tobytes()).hexdigest() == golden["best_ranks_sha256"]


def _small(name, iterations=131072):
    d = WORKLOADS[name]()
    d["iterations"] = iterations
    d["chunk_size"] = min(iterations, 100000)
    return d


BATCH_MIX = ["fcc108_ternary", "bcc128_quaternary", "hcp128_binary", "rocksalt216_split",
             "fcc256_quinary", "fcc32_3shells"]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_matches_single_job_runs(gpu, use_graph):
    """uneven jobs (108/128/128/216/256/32 sites, 2-5 species, 3-7 shells) in shared
    persistent launches return exactly the packs of their single-job runs"""
    print(f"[batch] graph={use_graph}", flush=True)
    alone = []
    for name in BATCH_MIX:
        job = Job(parse_config(_small(name)))
        job.set_kernel(2)
        job.run()
        alone.append(_pack_signature(job.results()))
        job.close()
    jobs = [Job(parse_config(_small(name))) for name in BATCH_MIX]
    for j in jobs:
        j.set_launch_size(32768)  # several persistent steps per job
    rep = S.batch_run(jobs, use_graph=use_graph)
    assert rep["candidates"] == 131072 * len(jobs)
    assert rep["launches"] >= 4
    for j, ref in zip(jobs, alone):
        assert j.timing()[2] == 131072
        assert _pack_signature(j.results()) == ref
        assert Rescorer(j.config).check_pack(j.results()) > 0
        j.close()


def test_batch_systematic_all_results(gpu):
    d = W.reference_test_16()
    job = Job(parse_config(d))
    job.set_launch_size(2000)
    S.batch_run([job])
    pack = job.results()
    assert pack.num_results() == 12870
    ranks = np.sort(pack.arrays(species=False)["rank"])
    assert np.array_equal(ranks, np.arange(1, 12871))


def test_server_results_alone_and_in_the_mix(gpu):
    """continuous batching: jobs admitted while others are resident produce the same packs
    as when run alone"""
    names = ["fcc256_quinary", "fcc32_3shells", "hcp128_binary", "fcc108_ternary",
             "rocksalt216_split", "bcc128_quaternary", "fcc64_ternary", "fcc500_quaternary",
             "fcc864_quinary"]
    its = [65536, 200000, 100000, 70000, 50000, 90000, 120000, 30000, 12000]
    alone = []
    for name, it in zip(names, its):
        job = Job(parse_config(_small(name, it)))
        job.set_kernel(2)
        job.run()
        alone.append(_pack_signature(job.results()))
        job.close()
    srv = S.Server(max_active=3, slice_candidates=16384)
    jobs = [Job(parse_config(_small(n, it))) for n, it in zip(names, its)]
    tickets = [srv.submit(jobs[0]), srv.submit(jobs[1])]
    srv.step()
    tickets += [srv.submit(j) for j in jobs[2:4]]
    srv.step()
    srv.step()
    tickets += [srv.submit(j) for j in jobs[4:7]]
    srv.step()
    tickets += [srv.submit(j) for j in jobs[7:]]
    srv.drain()
    for t, j, ref in zip(tickets, jobs, alone):
        st = srv.poll(t)
        assert st["done"] and st["t_admit"] >= st["t_submit"] and st["t_done"] >= st["t_admit"]
        assert _pack_signature(j.results()) == ref
        assert Rescorer(j.config).check_pack(j.results()) > 0
    srv.close()


def test_autotune_picks_a_variant(gpu):
    job = Job(parse_config(_small("fcc256_quinary")))
    kernel, rate = job.autotune()
    assert kernel in (1, 2, 3, 4) and rate > 0
    job.run()
    assert job.timing()[2] == 131072


@pytest.mark.parametrize("name", ["fcc32_3shells", "fcc64_ternary", "fcc108_ternary",
                                  "fcc256_quinary", "fcc864_quinary"])
def test_dense_launch_shapes_same_packs(gpu, name):
    """every alternative dense launch shape of the cell's tile bound, fewer candidate passes
    per wavefront round and other launch sizes return the default shape's pack, in single-job
    runs and side by side in one batched run; the pack is re-scored by the oracle"""
    its = 20_000 if name == "fcc864_quinary" else 100_000
    d = _small(name, its)
    ref_job = Job(parse_config(d))
    ref_job.set_kernel(3)
    ref_job.run()
    ref = _pack_signature(ref_job.results())
    assert Rescorer(ref_job.config).check_pack(ref_job.results()) > 0
    n = ref_job.dense_launch()["num_shapes"]
    assert n >= 2, "every tile bound of these cells has an alternative shape"
    ref_job.close()
    settings = [(a, 0, 0) for a in range(1, n)] + [(0, 1, 0), (0, 2, 7_001), (n - 1, 1, 30_000)]
    jobs = []
    for shape, passes, launch in settings:
        print(f"[shapes] {name} shape={shape} passes={passes} launch={launch}", flush=True)
        job = Job(parse_config(d))
        job.set_kernel(3)
        job.set_dense_launch(shape, passes, launch)
        assert job.dense_launch()["shape"] == shape
        job.run()
        assert job.timing()[2] == its
        assert _pack_signature(job.results()) == ref, (shape, passes, launch)
        job.close()
        b = Job(parse_config(d))
        b.set_kernel(3)
        b.set_dense_launch(shape, passes, 0)
        b.set_launch_size(its // 3 + 1)
        jobs.append(b)
    S.batch_run(jobs)
    for (shape, passes, _), b in zip(settings, jobs):
        assert _pack_signature(b.results()) == ref, ("batched", shape, passes)
        b.close()


def test_autotune_chooses_launch_parameters(gpu):
    """the autotuner keeps a variant, a dense shape, a pass count and a launch size; a run
    with them returns the untuned run's pack"""
    d = _small("fcc256_quinary", 3_000_000)
    ref_job = Job(parse_config(d))
    ref_job.run()
    ref = _pack_signature(ref_job.results())
    ref_job.close()
    job = Job(parse_config(d))
    kernel, rate = job.autotune()
    tuned = job.dense_launch()
    print(f"[autotune] kernel={kernel} rate={rate:.3g} {job.kernel_shape()} {tuned}", flush=True)
    assert kernel in (1, 2, 3, 4) and rate > 0
    assert 0 <= tuned["shape"] < tuned["num_shapes"]
    job.run()
    assert job.timing()[2] == 3_000_000
    assert _pack_signature(job.results()) == ref
    job.close()


@pytest.mark.parametrize("name", ["fcc256_quinary", "fcc864_quinary"])
def test_large_cell_sparse_vs_dense(gpu, name):
    """4096 uniformly random candidates scored by the LDS pair-list kernel and by the int8
    one-hot x int8 shell-adjacency MFMA contraction: bit-exact bond counts and objectives,
    and the oracle on a sample"""
    print(f"[sparse-vs-dense] {name}", flush=True)
    cfg = parse_config(WORKLOADS[name]())
    job = Job(cfg)
    ref = oracle_inputs(cfg, job)
    rng = np.random.default_rng(4096)
    cands = random_candidates(ref["packed"], ref["bounds"], 4096, rng)
    b_lane, o_lane, s_lane = job.score(cands, 0, kernel=2)
    b_dense, o_dense, s_dense = job.score(cands, 0, kernel=3)
    assert np.array_equal(b_lane, b_dense)
    assert np.array_equal(o_lane.view(np.uint64), o_dense.view(np.uint64))
    assert np.array_equal(s_lane, s_dense)
    for b in rng.choice(4096, 32, replace=False):
        rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
        assert np.array_equal(b_dense[b].astype(np.uint64), rb)
    job.close()


@pytest.mark.parametrize("name", ["fcc256_quinary", "fcc864_quinary"])
def test_score_dense_shell_split_consecutive(gpu, name):
    """explicit scoring split by shell (batches too small to fill the device): consecutive
    calls of different sizes on one job share its scratch and per-column counters; each call
    equals the pair-list kernel bit for bit"""
    print(f"[shell-split-consecutive] {name}", flush=True)
    cfg = parse_config(WORKLOADS[name]())
    job = Job(cfg)
    ref = oracle_inputs(cfg, job)
    rng = np.random.default_rng(77)
    cands = random_candidates(ref["packed"], ref["bounds"], 4096, rng)
    b_w, o_w, s_w = job.score(cands, 0, kernel=1)
    for B in (4096, 1, 100, 2000, 4096, 37):
        b_d, o_d, s_d = job.score(cands[:B], 0, kernel=3)
        assert np.array_equal(b_d, b_w[:B]), B
        assert np.array_equal(o_d.view(np.uint64), o_w[:B].view(np.uint64)), B
        assert np.array_equal(s_d, s_w[:B]), B
    job.close()


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_dense_matches_single_job_runs(gpu, use_graph):
    """the persistent dense form (jobs of one tile bound share a launch) returns the packs of
    the single-job dense runs, which equal the lane kernel's"""
    print(f"[batch-dense] graph={use_graph}", flush=True)
    names = ["fcc256_quinary", "fcc108_ternary", "bcc128_quaternary", "hcp128_binary",
             "fcc32_3shells"]
    alone = []
    for name in names:
        job = Job(parse_config(_small(name)))
        job.set_kernel(2)
        job.run()
        alone.append(_pack_signature(job.results()))
        job.close()
    jobs = [Job(parse_config(_small(name))) for name in names]
    for j in jobs:
        j.set_kernel(3)
        j.set_launch_size(32768)
    rep = S.batch_run(jobs, use_graph=use_graph)
    assert rep["candidates"] == 131072 * len(jobs)
    for j, ref in zip(jobs, alone):
        assert j.timing()[2] == 131072
        assert _pack_signature(j.results()) == ref
        assert Rescorer(j.config).check_pack(j.results()) > 0
        j.close()


@pytest.mark.parametrize("name", PAIRED)
@pytest.mark.parametrize("iterations", [1, 100, 241, 30_000])
def test_paired_matches_lane_kernel(gpu, name, iterations):
    """the paired form (MFMA and helper wavefronts, block batches of four pair batches)
    returns the lane kernel's pack, also for ranges that leave pairs or whole blocks idle"""
    print(f"[paired] {name} its={iterations}", flush=True)
    packs = []
    for kernel in (2, 4):
        job = Job(parse_config(_small(name, iterations)))
        job.set_kernel(kernel)
        job.run()
        assert job.timing()[2] == iterations
        packs.append(_pack_signature(job.results()))
        job.close()
    assert packs[0] == packs[1]


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_paired_matches_single_job_runs(gpu, use_graph):
    """the persistent paired form groups jobs by tile bound and species count (one LDS
    layout per launch); two seeds of the 256-site job share a group, jobs of other shapes
    run in groups of their own, and every pack equals the lane kernel's"""
    print(f"[batch-paired] graph={use_graph}", flush=True)
    cfgs = []
    for name in PAIRED:
        cfgs.append(_small(name))
    extra = _small("fcc256_quinary")
    extra["seed"] = 4242
    cfgs.append(extra)
    alone = []
    for d in cfgs:
        job = Job(parse_config(d))
        job.set_kernel(2)
        job.run()
        alone.append(_pack_signature(job.results()))
        job.close()
    jobs = [Job(parse_config(d)) for d in cfgs]
    for j in jobs:
        j.set_kernel(4)
        j.set_launch_size(32768)
    rep = S.batch_run(jobs, use_graph=use_graph)
    assert rep["candidates"] == 131072 * len(jobs)
    for j, ref in zip(jobs, alone):
        assert j.timing()[2] == 131072
        assert _pack_signature(j.results()) == ref
        j.close()


def test_paired_rejects_systematic_and_large_cells(gpu):
    d = W.reference_test_16()
    del d["pair_weights"]
    job = Job(parse_config(d))
    with pytest.raises(ValueError):
        job.set_kernel(4)
        job.run()
    job.close()
    job = Job(parse_config(_small("fcc864_quinary", 1000)))
    with pytest.raises(ValueError):
        job.set_kernel(4)
        job.run()
    job.close()


@pytest.mark.parametrize("name", ["fcc256_quinary", "fcc108_ternary", "bcc128_quaternary",
                                  "hcp128_binary", "fcc32_3shells", "fcc864_quinary"])
def test_pruned_dense_matches_lane_kernel(gpu, name):
    """early exit of hopeless passes (dense kernel, non-negative pair weights) keeps the
    result pack bit-identical and still counts every candidate"""
    print(f"[pruned] {name}", flush=True)
    its = 200_000 if name != "fcc864_quinary" else 20_000
    ref = Job(parse_config(_small(name, its)))
    ref.set_kernel(2)
    ref.run()
    want = _pack_signature(ref.results())
    ref.close()
    job = Job(parse_config(_small(name, its)))
    job.set_kernel(3)
    job.set_pruning(True)
    job.set_launch_size(32768)
    job.run()
    assert job.timing()[2] == its
    assert _pack_signature(job.results()) == want
    job.close()


def test_pruned_batch_and_systematic(gpu):
    names = ["fcc256_quinary", "fcc108_ternary", "hcp128_binary"]
    alone = []
    for name in names:
        j = Job(parse_config(_small(name)))
        j.set_kernel(2)
        j.run()
        alone.append(_pack_signature(j.results()))
        j.close()
    jobs = [Job(parse_config(_small(name))) for name in names]
    for j in jobs:
        j.set_kernel(3)
        j.set_pruning(True)
        j.set_launch_size(32768)
    S.batch_run(jobs, use_graph=True)
    for j, want in zip(jobs, alone):
        assert _pack_signature(j.results()) == want
        j.close()
    d = W.reference_test_16()
    del d["pair_weights"]
    packs = []
    for prune in (False, True):
        j = Job(parse_config(d))
        j.set_kernel(3)
        j.set_pruning(prune)
        j.set_launch_size(1000)
        j.run()
        packs.append(_pack_signature(j.results()))
        j.close()
    assert packs[0] == packs[1]


def _cell(name):
    """cells that exercise the remaining dense shapes: tile bounds 6 (192 sites), 12 (384),
    14 (432 sites: 13.5 tiles, a partial last tile), 16 (500), 20 (600 sites: 19 tiles
    under bound 20, the stream's grid positions past the problem hold zero tiles), 24 (768)
    and 32 (1024 sites, the largest), 16 species (the dense kernel's maximum, beyond the
    lane kernel) and 96 sites (3 tiles under bound 4)"""
    metals16 = ["Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Zr", "Nb", "Mo", "Ru",
                "Rh", "Pd", "Ag"]
    if name == "fcc500_quinary":
        comp = {"Fe": 100, "Co": 100, "Ni": 100, "Cr": 100, "Mn": 100}
        return W.fcc(4.05, (5, 5, 5), comp, 4, seed=71, thread_config=1)
    if name == "fcc1024_ternary":
        comp = {"Fe": 342, "Co": 341, "Ni": 341}
        return W.fcc(4.05, (8, 8, 4), comp, 3, seed=72, thread_config=1)
    if name == "fcc256_16species":
        comp = {m: 16 for m in metals16}
        return W.fcc(4.05, (4, 4, 4), comp, 3, seed=73, thread_config=1)
    if name == "fcc96_quaternary":
        comp = {"Fe": 24, "Co": 24, "Ni": 24, "Cr": 24}
        return W.fcc(4.05, (2, 3, 4), comp, 4, seed=74, thread_config=1)
    if name == "fcc192_quinary":
        comp = {"Fe": 39, "Co": 39, "Ni": 38, "Cr": 38, "Mn": 38}
        return W.fcc(4.05, (4, 4, 3), comp, 5, seed=75, thread_config=1)
    if name == "fcc384_quaternary":
        comp = {"Fe": 96, "Co": 96, "Ni": 96, "Cr": 96}
        return W.fcc(4.05, (4, 4, 6), comp, 4, seed=76, thread_config=1)
    if name == "fcc432_ternary":
        comp = {"Fe": 144, "Co": 144, "Ni": 144}
        return W.fcc(4.05, (3, 4, 9), comp, 5, seed=77, thread_config=1)
    if name == "fcc600_quinary":
        comp = {"Fe": 120, "Co": 120, "Ni": 120, "Cr": 120, "Mn": 120}
        return W.fcc(4.05, (5, 5, 6), comp, 4, seed=78, thread_config=1)
    if name == "fcc768_ternary":
        comp = {"Fe": 256, "Co": 256, "Ni": 256}
        return W.fcc(4.05, (4, 8, 6), comp, 3, seed=79, thread_config=1)
    raise KeyError(name)


@pytest.mark.parametrize("name,ref_kernel", [("fcc500_quinary", 2), ("fcc1024_ternary", 2),
                                             ("fcc256_16species", 1), ("fcc96_quaternary", 2),
                                             ("fcc192_quinary", 2), ("fcc384_quaternary", 2),
                                             ("fcc432_ternary", 2), ("fcc600_quinary", 2),
                                             ("fcc768_ternary", 2)])
def test_dense_shapes_match_sparse(gpu, name, ref_kernel):
    """explicit scoring of random candidates (bit-exact bonds, objectives and SRO against the
    sparse kernel, bonds against the oracle on a sample) and a short MC run whose pack equals
    the sparse kernel's"""
    print(f"[dense-shapes] {name}", flush=True)
    cfg = parse_config(_cell(name))
    assert not isinstance(cfg, S.ParseError), cfg
    job = Job(cfg)
    ref = oracle_inputs(cfg, job)
    rng = np.random.default_rng(11)
    cands = random_candidates(ref["packed"], ref["bounds"], 512, rng)
    b_ref, o_ref, s_ref = job.score(cands, 0, kernel=ref_kernel)
    b_d, o_d, s_d = job.score(cands, 0, kernel=3)
    assert np.array_equal(b_ref, b_d)
    assert np.array_equal(o_ref.view(np.uint64), o_d.view(np.uint64))
    assert np.array_equal(s_ref, s_d)
    for b in rng.choice(512, 8, replace=False):
        rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
        assert np.array_equal(b_d[b].astype(np.uint64), rb)
    job.close()
    packs = []
    for kernel in (ref_kernel, 3):
        d = _cell(name)
        d["iterations"] = 20_000
        d["chunk_size"] = 20_000
        j = Job(parse_config(d))
        j.set_kernel(kernel)
        j.run()
        assert j.timing()[2] == 20_000
        packs.append(_pack_signature(j.results()))
        j.close()
    assert packs[0] == packs[1]


def _mid_cell(name, iterations):
    d = WORKLOADS[name]() if name in WORKLOADS else _cell(name)
    d["iterations"] = iterations
    d["chunk_size"] = min(iterations, 100000)
    return d


MID_BOUNDS = ["fcc192_quinary", "fcc384_quaternary", "fcc432_ternary", "fcc600_quinary",
              "fcc768_ternary", "fcc864_quinary"]  # tile bounds 6, 12, 14, 20, 24, 27


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_dense_mid_bounds_match_lane_kernel(gpu, use_graph):
    """the persistent dense form over cells of six tile bounds (one launch group per bound,
    the last launch of each job partial) returns the lane kernel's packs"""
    print(f"[batch-dense-mid] graph={use_graph}", flush=True)
    its = 20_000
    alone = []
    for name in MID_BOUNDS:
        job = Job(parse_config(_mid_cell(name, its)))
        job.set_kernel(2)
        job.run()
        alone.append(_pack_signature(job.results()))
        job.close()
    jobs = [Job(parse_config(_mid_cell(name, its))) for name in MID_BOUNDS]
    for j in jobs:
        j.set_kernel(3)
        j.set_launch_size(8192)
    shapes = [j.kernel_shape() for j in jobs]
    assert [s.split(",")[0] for s in shapes] == ["dense<6", "dense<12", "dense<14", "dense<20",
                                                "dense<24", "dense<27"]
    rep = S.batch_run(jobs, use_graph=use_graph)
    assert rep["candidates"] == its * len(jobs)
    for j, ref in zip(jobs, alone):
        assert j.timing()[2] == its
        assert _pack_signature(j.results()) == ref
        j.close()


@pytest.mark.parametrize("iterations", [1, 241, 30_000])
def test_paired_bound6_matches_lane_kernel(gpu, iterations):
    """the paired form at tile bound 6 (192 sites, ring depth 7) returns the lane kernel's
    pack"""
    print(f"[paired-6] its={iterations}", flush=True)
    packs = []
    for kernel in (2, 4):
        job = Job(parse_config(_mid_cell("fcc192_quinary", iterations)))
        job.set_kernel(kernel)
        if kernel == 4:
            assert job.kernel_shape() == "paired<6,2,7>"
        job.run()
        assert job.timing()[2] == iterations
        packs.append(_pack_signature(job.results()))
        job.close()
    assert packs[0] == packs[1]


_SC_SPECIES = ["Ti", "V", "Cr", "Mn", "Fe", "Co", "Ni", "Cu", "Zn", "Zr", "Nb", "Mo", "Ru",
               "Rh", "Pd", "Ag"]


def _sc_cell(dims, K, shells, seed):
    """simple cubic supercell of any site count: K species as even as the count allows"""
    n = int(np.prod(dims))
    base, extra = divmod(n, K)
    comp = {_SC_SPECIES[i]: base + (1 if i < extra else 0) for i in range(K)}
    return dict(structure=dict(lattice=np.diag([3.0, 3.0, 3.0]).tolist(), coords=[[0, 0, 0]],
                               species=["Al"], supercell=list(dims)),
                composition=comp,
                shell_weights={str(s): 1.0 / s for s in range(1, shells + 1)},
                iterations=4096, chunk_size=4096, seed=seed, thread_config=1)


@pytest.mark.parametrize("dims,K,shells,bound", [
    ((3, 3, 3), 2, 3, 1), ((3, 3, 4), 2, 3, 2), ((3, 4, 6), 3, 4, 4), ((4, 5, 8), 4, 3, 6),
    ((5, 5, 9), 5, 4, 8), ((5, 7, 10), 6, 3, 12), ((6, 7, 10), 7, 3, 14), ((6, 8, 10), 3, 4, 16),
    ((7, 8, 11), 4, 3, 20), ((8, 9, 10), 2, 3, 24), ((7, 11, 11), 4, 3, 27),
    ((8, 10, 11), 5, 3, 32), ((9, 10, 11), 3, 3, 32), ((5, 7, 10), 11, 3, 12),
    ((7, 8, 11), 16, 2, 20), ((8, 10, 11), 13, 2, 32)])
def test_dense_odd_cells_match_pair_list(gpu, dims, K, shells, bound):
    """cells whose site count is no multiple of 32 (a partial last tile, and tiles past the
    problem up to the bound) at every tile bound: explicit scoring by the dense kernel equals
    the pair-list kernel bit for bit (bonds, objectives, SRO) and the oracle on a sample,
    and a short Monte Carlo run returns the lane kernel's pack (the wave kernel's beyond the
    lane kernel's 8 species)"""
    n = int(np.prod(dims))
    print(f"[dense-odd] {n} sites K={K} S={shells}", flush=True)
    cfg = parse_config(_sc_cell(dims, K, shells, n))
    assert not isinstance(cfg, S.ParseError), cfg
    job = Job(cfg)
    assert job.kernel_shape().startswith(f"dense<{bound},")
    ref = oracle_inputs(cfg, job)
    rng = np.random.default_rng(n)
    B = 128 if n < 500 else 32
    cands = random_candidates(ref["packed"], ref["bounds"], B, rng)
    b_w, o_w, s_w = job.score(cands, 0, kernel=1)
    b_d, o_d, s_d = job.score(cands, 0, kernel=3)
    assert np.array_equal(b_w, b_d)
    assert np.array_equal(o_w.view(np.uint64), o_d.view(np.uint64))
    assert np.array_equal(s_w, s_d)
    for b in rng.choice(B, 4, replace=False):
        rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
        assert np.array_equal(b_d[b].astype(np.uint64), rb)
    job.close()
    packs = []
    for kernel in (2 if K <= 8 else 1, 3):
        j = Job(parse_config(_sc_cell(dims, K, shells, n)))
        j.set_kernel(kernel)
        j.run()
        assert j.timing()[2] == 4096
        packs.append(_pack_signature(j.results()))
        j.close()
    assert packs[0] == packs[1]


@pytest.mark.parametrize("name,B", [("fcc108_ternary", 49152), ("fcc256_quinary", 32768)])
def test_score_dense_large_batch_tile_groups(gpu, name, B):
    """explicit scoring of a batch that fills the device at one pass per wavefront runs the
    shape's tile groups per wavefront; a small batch runs one group: both equal the pair-list
    kernel bit for bit, and each other on the candidates they share"""
    print(f"[score-large-batch] {name} B={B}", flush=True)
    cfg = parse_config(WORKLOADS[name]())
    job = Job(cfg)
    ref = oracle_inputs(cfg, job)
    rng = np.random.default_rng(B)
    cands = random_candidates(ref["packed"], ref["bounds"], B, rng)
    b_w, o_w, s_w = job.score(cands, 0, kernel=1)
    b_d, o_d, s_d = job.score(cands, 0, kernel=3)
    assert np.array_equal(b_w, b_d)
    assert np.array_equal(o_w.view(np.uint64), o_d.view(np.uint64))
    assert np.array_equal(s_w, s_d)
    b_s, o_s, s_s = job.score(cands[:4096], 0, kernel=3)
    assert np.array_equal(b_s, b_d[:4096])
    assert np.array_equal(o_s.view(np.uint64), o_d[:4096].view(np.uint64))
    assert np.array_equal(s_s, s_d[:4096])
    for b in rng.choice(B, 8, replace=False):
        rb = O.count_bonds(ref["pairs"], cands[b], ref["S"], ref["K"])
        assert np.array_equal(b_d[b].astype(np.uint64), rb)
    job.close()
