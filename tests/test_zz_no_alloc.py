"""sqsg_run_noalloc (no device allocations; its scratch is a fixed device array of the
library) against sqsg_run: identical result packs on consecutive calls, the second with a
smaller problem, and when a launch's records overflow the array."""
import ctypes as C

import numpy as np
import pytest

import workloads as W
import sqsgenerator_amd._lib as L
from sqsgenerator_amd.core import Job, parse_config

pytestmark = pytest.mark.gpu


def _signature(pack):
    arr = pack.arrays(species=True, ranks=False)
    return (pack.level_objectives.view(np.uint64).tolist(), pack.level_start.tolist(),
            arr["objective"].view(np.uint64).tolist(), arr["candidate"].tolist(),
            arr["species"].tobytes())


def _noalloc(job):
    rc = L.lib().sqsg_run_noalloc(job.handle, L.PROGRESS_FN(), None)
    assert rc == 0, L.lib().sqsg_last_error()


def test_noalloc_matches_run_on_consecutive_calls(gpu):
    cases = [W.quinary_256(iterations=300_000, seed=31),    # dense MFMA kernel, 256 sites
             dict(W.plain_mc_32(iterations=50_000, seed=32), chunk_size=10_000)]
    jobs = [Job(parse_config(d)) for d in cases]
    refs = []
    for job in jobs:
        job.run()  # uploads the inputs and gives the reference pack
        pack = job.results()
        refs.append((_signature(pack), pack.sro_array()))
    for job, (sig, sro) in zip(jobs, refs):
        print(f"[noalloc] {job.kernel_shape()} {job.config.iterations}", flush=True)
        _noalloc(job)
        pack = job.results()
        assert pack.statistics["finished"] == job.config.iterations
        assert _signature(pack) == sig
        assert np.array_equal(pack.sro_array(), sro)
        job.close()


def test_noalloc_overflowing_scratch_and_systematic(gpu):
    """keep above the candidate count records every candidate: a launch of 65,536 records of
    256 sites exceeds the 16 MiB array, so launches shrink until they fit; a systematic job
    (records not bounded by the wavefront filter) follows"""
    d = W.quinary_256(iterations=200_000, seed=33)
    d["keep"] = 1_000_000
    sys16 = W.reference_test_16()
    del sys16["pair_weights"]
    sys16.update(keep=3, chunk_size=1000)
    for cfg in (d, sys16):
        job = Job(parse_config(cfg))
        job.run()
        sig = _signature(job.results())
        _noalloc(job)
        pack = job.results()
        assert _signature(pack) == sig
        job.close()


def test_noalloc_needs_resident_inputs(gpu):
    job = Job(parse_config(dict(W.plain_mc_32(iterations=1000), chunk_size=1000)))
    rc = L.lib().sqsg_run_noalloc(job.handle, L.PROGRESS_FN(), None)
    assert rc == L.ERR_INVALID_ARGUMENT
    assert b"resident" in L.lib().sqsg_last_error()
    job.close()
