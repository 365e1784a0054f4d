"""Benchmark: candidates scored per second on the 256-site quinary FCC cell, 7 shells.

One step = one device launch that generates, scores (bond histogram + FP64 objective) and
tie-filters a batch of MC candidates of the FCC 4x4x4 Fe/Co/Ni/Cr/Mn (51/51/51/51/52) job.
Prints one JSON line: throughput, the dominant kernel's roofline fraction and the CPU
baseline measured in the same run.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "candidates scored/sec (256-site 5-species FCC, 7 shells); job p99 latency"
HBM_PEAK_GBS = 8000.0


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(job, cfg, budget_s: float) -> dict:
    """The oracle's restatement of the reference worker loop (sqs.h:348-455), multithreaded
    with the shared RNG state, on a bounded sample of the same workload."""
    import numpy as np
    from oracle import oracle as O

    info = job.problem(0)
    prob = O.Problem(info["pairs"].astype(np.uint64), info["species_packed"],
                     np.array(info["bounds"]).ravel(), info["prefactors"], info["pair_weights"],
                     info["target"], seed=7)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    # calibrate on a short single-thread run, then size the sample to the budget
    st, _ = O.run([prob], 2000, chunk_size=2000, threads=1, collect=False)
    rate1 = st.finished / max(st.seconds, 1e-9)
    n = int(rate1 * threads * budget_s)
    n = max(threads * 1000, n)
    chunk = max(1000, n // (threads * 8))
    st, _ = O.run([prob], n, chunk_size=chunk, threads=threads, shared_rng=True, collect=False)
    return {"value": st.finished / st.seconds, "unit": "candidates/s", "cores": threads,
            "kind": "port",
            "sample": f"{st.finished} MC candidates of the same 256-site job, "
                      f"{threads} threads, shared RNG state, {st.seconds:.1f} s "
                      f"on {cpu_model()}"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--candidates-per-step", type=int, default=1 << 20)
    ap.add_argument("--kernel", default="auto", choices=["auto", "wave", "lane"])
    ap.add_argument("--plain-steps", type=int, default=3,
                    help="launches of the plain wave-per-candidate kernel timed for comparison")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration_copy_bytes": n, "launches": 3}))
        return 0
    if args.gpus != 1:
        print("bench.py measures one GPU only: candidates are independent, so N GPUs would "
              "each run disjoint RNG streams / rank ranges and merge their tie sets on the "
              "host, but multi-GPU runs are not part of this build", file=sys.stderr)
        return 2

    import numpy as np

    import sqsgenerator_amd._lib as L
    import workloads as W
    from sqsgenerator_amd.core import Job, parse_config

    per_step = args.candidates_per_step
    kernel = {"auto": L.KERNEL_AUTO, "wave": L.KERNEL_SPARSE_WAVE,
              "lane": L.KERNEL_SPARSE_LANE}[args.kernel]

    def make(iterations, k=kernel):
        cfg = parse_config(W.quinary_256(iterations=iterations, seed=7))
        job = Job(cfg)
        job.set_kernel(k)
        job.set_launch_size(per_step)
        return cfg, job

    plain = None
    if args.plain_steps:
        _, pjob = make(args.plain_steps * per_step, L.KERNEL_SPARSE_WAVE)
        pjob.run()
        pms, _, pn = pjob.timing()
        plain = pn / (pms / 1e3)
        pjob.close()

    # warmup: W launches of the same job
    if args.warmup:
        _, wjob = make(args.warmup * per_step)
        wjob.run()
        wjob.close()
    cfg, job = make(args.steps * per_step)
    info = job.problem(0)
    job.score(info["species_packed"][None, :])   # uploads the job: inputs resident in HBM
    L.check(L.lib().sqsg_device_synchronize())
    t0 = time.perf_counter()
    job.run()
    L.check(L.lib().sqsg_device_synchronize())
    wall = time.perf_counter() - t0
    kernel_ms, launches, scored = job.timing()
    assert scored == args.steps * per_step, (scored, args.steps * per_step)
    value = scored / wall
    kernel_name = {L.KERNEL_SPARSE_WAVE: "run_wave_kernel"}.get(kernel, "run_lane_kernel<5>")
    avg_launch_s = kernel_ms / 1e3 / launches
    P, N = info["num_pairs"], info["num_sites"]
    bytes_per_cand = 6 * P + N + 8
    achieved = bytes_per_cand * per_step / avg_launch_s / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "candidates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32+f64", "data": "synthetic",
        "config": {"workload": "fcc4x4x4-256site-quinary-7shell-mc", "sites": N,
                   "species": info["num_species"], "shells": info["num_shells"], "pairs": P,
                   "candidates_per_step": per_step, "kernel": args.kernel,
                   "kernel_candidates_per_s": scored / (kernel_ms / 1e3),
                   "plain_kernel_candidates_per_s": plain,
                   "speedup_vs_plain": (scored / (kernel_ms / 1e3)) / plain if plain else None},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": kernel_name, "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_bytes_per_candidate": bytes_per_cand},
    }
    job.close()
    if not args.no_cpu:
        _, cjob = make(per_step)
        out["cpu_baseline"] = cpu_baseline(cjob, cfg, args.cpu_budget)
        cjob.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
