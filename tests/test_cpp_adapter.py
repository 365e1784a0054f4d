"""A C++ caller of the C ABI (tests/cpp/adapter.cpp), compiled with the host compiler against
include/sqsgen_amd.h and linked to the built library: the adapter a maintainer adds on the
reference side of run_optimization (sqs.h:625-634).  The host checks need no device; the GPU
tests compare the packs it rebuilds with the Python binding's for the same configuration."""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import workloads as W
from gpu_oracle import Rescorer
from sqsgenerator_amd.core import Job, parse_config

ROOT = Path(__file__).resolve().parents[1]
LIB_DIR = ROOT / "sqsgenerator_amd" / "lib"


@pytest.fixture(scope="module")
def adapter(tmp_path_factory):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    if not (LIB_DIR / "libsqsgen_amd.so").exists():
        pytest.fail("libsqsgen_amd.so is not built")
    exe = tmp_path_factory.mktemp("adapter") / "adapter"
    cmd = [cxx, "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-I", str(ROOT / "include"),
           str(ROOT / "tests" / "cpp" / "adapter.cpp"), "-L", str(LIB_DIR), "-lsqsgen_amd",
           f"-Wl,-rpath,{LIB_DIR}", "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, *args, timeout=120):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


def test_adapter_host_checks(adapter):
    """geometry, ranks, problem construction and the status -> exception mapping"""
    assert _run(adapter, "host") == {"host_checks_failed": 0}


def _binding_pack(iterations, stop_after=None):
    cfg = parse_config(W.plain_mc_32(iterations=iterations, seed=42))
    job = Job(cfg)
    job.set_launch_size(50_000)
    calls = []

    def cb(ctx):
        calls.append(ctx.statistics["finished"])
        if stop_after is not None and len(calls) >= stop_after:
            ctx.stop()

    job.run(cb)
    return cfg, job, job.results()


def _same_pack(doc, pack):
    assert len(doc["levels"]) == len(pack)
    for l, lv in enumerate(doc["levels"]):
        assert float.fromhex(lv["objective"]) == pack.level_objectives[l]
        a, b = int(pack.level_start[l]), int(pack.level_start[l + 1])
        arr = pack.arrays(a, b - a, species=True, ranks=True)
        sro = pack.sro_array(a, b - a)
        assert len(lv["results"]) == b - a
        for k, res in enumerate(lv["results"]):
            assert float.fromhex(res["objective"]) == arr["objective"][k]
            assert np.array_equal(np.array(res["species"], np.uint8), arr["species"][k])
            assert int(res["rank"]) == int(arr["rank"][k])
            assert [float.fromhex(x) for x in res["sro"]] == sro[k].tolist()


@pytest.mark.gpu
def test_adapter_run_matches_binding(gpu, adapter):
    """the C++ caller's pack equals the binding's, level by level and result by result,
    and every result re-scores to its objective on the oracle"""
    doc = _run(adapter, "run", 200_000)
    assert doc["finished"] == 200_000 and doc["launches"] == 4
    cfg, job, pack = _binding_pack(200_000)
    _same_pack(doc, pack)
    assert Rescorer(cfg).check_pack(pack) == pack.num_results() > 0
    job.close()


@pytest.mark.gpu
def test_adapter_stop_returns_partial_pack(gpu, adapter):
    """the callback's stop after the first launch ends the C++ caller's run with 50,000 of
    200,000 candidates and the pack of those candidates"""
    doc = _run(adapter, "stop", 200_000)
    assert doc["finished"] == 50_000 and doc["launches"] == 1
    cfg, job, pack = _binding_pack(200_000, stop_after=1)
    assert pack.statistics["finished"] == 50_000
    _same_pack(doc, pack)
    assert Rescorer(cfg).check_pack(pack) == pack.num_results() > 0
    job.close()
