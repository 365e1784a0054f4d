"""GPU parity of the result collection (core/results.h:42-99, sqs.h:384-441,461-466).

The device path's result packs are compared, level by level and configuration by
configuration, with the oracle's driver run on one thread: the reference's blocks walked in
order (BS::thread_pool's near-equal split, each interact block starting from the sorted
configuration), insertion at or below the keep-th distinct inserted objective, levels never
evicted, at most max_results_per_objective + 1 configurations per level and the stall of an
interact block after such a skip.  Systematic runs enumerate the reference's own candidate
stream; Monte Carlo runs use the device's counter-based candidates, which the oracle
regenerates (counter_stream).
"""
import numpy as np
import pytest

import workloads as W
import sqsgenerator_amd as S
from gpu_oracle import Rescorer, assert_packs_equal, device_pack, oracle_pack
from sqsgenerator_amd.core import Job, batch_run, parse_config

pytestmark = pytest.mark.gpu


def _sys16(**kw):
    """the 12870-candidate space of python/tests/test_core.py:27-42 with the default pair
    weights (levels of a few ties each)"""
    d = W.reference_test_16(**kw)
    del d["pair_weights"]
    return d


def _run(d, kernel=0, launch=None, callback=None):
    cfg = parse_config(d)
    job = Job(cfg)
    if kernel:
        job.set_kernel(kernel)
    if launch:
        job.set_launch_size(launch)
    job.run(callback)
    return cfg, job, job.results()


@pytest.mark.parametrize("chunk", [1000, 12870])
@pytest.mark.parametrize("max_results", [None, 1, 2])
@pytest.mark.parametrize("keep", [1, 3])
@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_systematic_pack_matches_oracle(gpu, kernel, keep, max_results, chunk):
    """the whole pack of the 16-site systematic search: the reference's own candidate
    stream, so every level and every configuration is pinned"""
    print(f"[sys16] kernel={kernel} keep={keep} max={max_results} chunk={chunk}", flush=True)
    d = _sys16(keep=keep, chunk_size=chunk)
    if max_results is not None:
        d["max_results_per_objective"] = max_results
    cfg, job, pack = _run(d, kernel, launch=2000 if chunk == 1000 else None)
    assert pack.statistics["finished"] == 12870
    assert_packs_equal(device_pack(pack), oracle_pack(cfg, job))
    Rescorer(cfg).check_pack(pack)
    if max_results is not None:
        for l in range(len(pack)):
            assert pack.level_start[l + 1] - pack.level_start[l] <= max_results + 1
    job.close()


def _split_uneven(seed=14):
    """rocksalt 3x3x3 in split mode with sublattices of different size and species count:
    the 108 cation sites hold Ti/Al/Zr, 54 of the anion sites N/B, the others stay N"""
    d = W.rocksalt_split_216(seed=seed)
    st = parse_config(d).structure()
    anions = [i for i, z in enumerate(st.species) if z == 7][:54]
    d["composition"] = [{"sites": "Ti", "Ti": 36, "Al": 36, "Zr": 36},
                        {"sites": anions, "N": 27, "B": 27}]
    d["target_objective"] = [0, 0]
    return d


MC_CASES = {
    # name: (config, iterations, chunk, kernels)
    "sys16_mc": (lambda: dict(_sys16(), iteration_mode="random", seed=5), 30001, 7000,
                 (1, 2, 3, 4)),
    "fcc32": (lambda: W.plain_mc_32(seed=11), 50001, 10000, (1, 2, 3, 4)),
    "fcc256": (lambda: W.quinary_256(seed=12), 20001, 6000, (2, 3)),
    "rocksalt_split": (lambda: W.rocksalt_split_216(seed=13), 5001, 1500, (1, 2, 3)),
    "split_uneven": (_split_uneven, 5001, 1500, (1, 2, 3)),
    # the dense kernel without the last species' column, windows of different neighbour
    # counts and fixed sites; and with every column, one window of mixed counts
    "perovskite_bx": (lambda: W.perovskite_bx(seed=22), 20001, 6000, (1, 2, 3)),
    "perovskite_mixed": (lambda: W.perovskite_bx(seed=23, mixed=True), 20001, 6000, (1, 2, 3)),
}


@pytest.mark.parametrize("max_results", [None, 1])
@pytest.mark.parametrize("keep", [1, 3])
@pytest.mark.parametrize("name", list(MC_CASES))
def test_mc_pack_matches_oracle(gpu, name, keep, max_results):
    """Monte Carlo packs equal the oracle driver's over the device's candidate stream, for
    every kernel (wave, lane, dense, paired) and with a launch size that splits blocks"""
    make, iterations, chunk, kernels = MC_CASES[name]
    d = make()
    d.update(iterations=iterations, chunk_size=chunk, keep=keep)
    if max_results is not None:
        d["max_results_per_objective"] = max_results
    ref = None
    for kernel in kernels:
        print(f"[mc] {name} kernel={kernel} keep={keep} max={max_results}", flush=True)
        cfg, job, pack = _run(d, kernel, launch=12288)
        assert pack.statistics["finished"] == iterations
        if ref is None:
            ref = oracle_pack(cfg, job)
            Rescorer(cfg).check_pack(pack)
        assert_packs_equal(device_pack(pack), ref)
        job.close()


def test_mc_keep_everything_pins_block_starts(gpu):
    """keep above the number of candidates inserts every distinct configuration, so the pack
    is the set of all candidates: the unshuffled candidates sit exactly at the block starts
    of the thread pool's split (5001 = 7 blocks of 715, 715, 715, 714, 714, 714, 714)"""
    d = dict(_sys16(), iteration_mode="random", seed=3, iterations=5001, chunk_size=700,
             keep=100000)
    for kernel in (1, 2, 3):
        cfg, job, pack = _run(d, kernel, launch=1024)
        ref = oracle_pack(cfg, job)
        assert_packs_equal(device_pack(pack), ref)
        sorted_conf = job.problem(0)["species_packed"].tobytes()
        starts = [0, 715, 1430, 2145, 2859, 3573, 4287]
        arr = pack.arrays(species=False, ranks=False)
        got = {int(arr["candidate"][r]) for r in range(pack.num_results())
               if pack.species_packed(r).tobytes() == sorted_conf}
        assert got == {0}  # the first occurrence; later block starts repeat it
        job.close()
    # the stream differs from the one with every candidate shuffled exactly there
    shuffled = oracle_pack(cfg, job_for(d), chunk=5001)
    assert set().union(*map(set, shuffled.values())) != set().union(*map(set, ref.values()))


def job_for(d):
    return Job(parse_config(d))


@pytest.mark.parametrize("kernel", [2, 3])
def test_stop_returns_partial_pack(gpu, kernel):
    """SqsCallbackContext.stop() after the second launch (sqs.h:385-396,447-451): the run
    ends with finished < iterations and the pack is the collection of the candidates scored,
    each re-scored by the oracle"""
    d = W.plain_mc_32(iterations=1_000_000, seed=21)
    calls = []

    def cb(ctx):
        calls.append(ctx.statistics["finished"])
        if len(calls) == 2:
            ctx.stop()

    cfg, job, pack = _run(d, kernel, launch=100_000, callback=cb)
    finished = pack.statistics["finished"]
    assert len(calls) == 2 and finished == calls[-1] == 200_000
    assert job.timing()[2] == finished < cfg.iterations
    # candidates [0, 200000) are the first two blocks of the run (chunk 100000)
    ref = oracle_pack(cfg, job, iterations=finished, chunk=100_000)
    assert_packs_equal(device_pack(pack), ref)
    assert Rescorer(cfg).check_pack(pack) > 0
    job.close()


def test_callback_exception_propagates(gpu):
    d = W.plain_mc_32(iterations=1_000_000, seed=22)

    def cb(ctx):
        raise RuntimeError("from the callback")

    with pytest.raises(RuntimeError, match="from the callback"):
        S.optimize(parse_config(d), callback=cb)


def test_batch_packs_match_oracle(gpu):
    """the batched persistent form over uneven jobs (configs (a)-(d) and a perovskite cell
    of another tile bound, at 1e5 candidates each, keep 2, one of them with
    max_results_per_objective) against the oracle's driver"""
    makers = [W.ternary_108, W.quaternary_bcc_128, W.hcp_128, W.rocksalt_split_216,
              W.perovskite_bx]
    cfgs, jobs = [], []
    for k, make in enumerate(makers):
        d = make(iterations=100_001)
        d.update(chunk_size=30_000, keep=2)
        if k == 2:
            d["max_results_per_objective"] = 1
        cfg = parse_config(d)
        cfgs.append(cfg)
        jobs.append(Job(cfg))
    batch_run(jobs)
    for cfg, job in zip(cfgs, jobs):
        print(f"[batch] {job.kernel_shape()}", flush=True)
        pack = job.results()
        assert pack.statistics["finished"] == cfg.iterations
        assert_packs_equal(device_pack(pack), oracle_pack(cfg, job))
        Rescorer(cfg).check_pack(pack)
        job.close()


def test_batch_abcd_full_budget_rescored(gpu):
    """configs (a)-(d) at 1e7 candidates each in one batched run: every reported structure
    of every level is re-scored by the oracle (objective bits, SRO cubes, ranks)"""
    makers = [W.ternary_108, W.quaternary_bcc_128, W.hcp_128, W.rocksalt_split_216]
    cfgs = [parse_config(make(iterations=10_000_000)) for make in makers]
    jobs = [Job(c) for c in cfgs]
    rep = batch_run(jobs)
    print(f"[abcd] {rep}", flush=True)
    for cfg, job in zip(cfgs, jobs):
        pack = job.results()
        assert pack.statistics["finished"] == 10_000_000
        assert job.timing()[2] == 10_000_000
        assert Rescorer(cfg).check_pack(pack) == pack.num_results() > 0
        job.close()
