"""Benchmark: candidates scored per second on the 256-site quinary FCC cell, 7 shells, plus
serving latency.

Workload: FCC 4x4x4 (a = 4.05), Fe/Co/Ni/Cr/Mn 51/51/51/51/52, 7 shells with weights 1/s,
uniform random MC candidates (counter-based streams), synthetic.  A candidate is one full
configuration -> shell x species^2 bond counts -> FP64 objective -> tie-keeping update.

  * plain: the single-job wavefront-per-candidate kernel, one launch per step;
  * batched (the reported value): the persistent, graph-captured multi-job form running
    four such jobs (different seeds) together with the kernel the job picks by default
    (the exact int8 MFMA contraction for this cell); one step = one persistent launch;
  * serving: Poisson arrivals of mixed SQS jobs on the continuous-batching server, p50/p99
    job latency at the stated arrival rate;
  * cpu_baseline: the oracle's restatement of the reference worker loop on the host cores.
Prints one JSON line.
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
INT8_PEAK_TOPS = 5000.0   # MI355X dense int8 MFMA, 5.0 POP/s
TRAFFIC_FILE = ROOT / "profiles" / "batch_dense_kernel_fcc256_traffic.json"


def dense_ops_per_candidate(N, K, S):
    """int8 ops of the dense one-hot contraction of one candidate, S (2 N^2 K + 2 N K^2)"""
    return S * (2 * N * N * K + 2 * N * K * K)


def executed_ops_per_candidate(shape, S):
    """int8 ops the dense kernel of launch shape `shape` ("dense<TT,G,W,D> columns C")
    executes per candidate: per shell and tile group, TT (TT + 1) / 2 GEMM1 and TT GEMM2
    32x32x32 products for 32 // C candidates"""
    head, _, cols = shape.partition(" columns ")
    TT = int(head[head.index("<") + 1:].split(",")[0])
    C = int(cols.split()[0].split("+")[0])
    return S * (TT * (TT + 1) // 2 + TT) * 2 * 32 ** 3 // (32 // C)


def measured_traffic(candidates_per_launch):
    """HBM bytes per launch from the committed rocprofv3 counter summary (FETCH_SIZE and
    WRITE_SIZE per dispatch, scaled by the copy-kernel calibration), if it matches this
    launch size"""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if t.get("candidates_per_launch") != candidates_per_launch:
        return None
    return t.get("hbm_bytes_per_launch")


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_share() -> dict:
    """Host threads the reference's default would use (std::thread::hardware_concurrency(),
    io/config/combined.h:105-107) and the cores this process may actually run on: its CPU
    affinity and the cgroup's CPU quota (threads beyond those only time-share the same
    cores)"""
    hw = os.cpu_count() or 1
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    usable = min(hw, affinity, int(quota) if quota else hw)
    return {"hardware_concurrency": hw, "affinity": affinity, "cgroup_cpu_quota": quota,
            "threads": max(1, usable)}


def cpu_baseline(job, budget_s: float) -> dict:
    """The oracle's restatement of the reference worker loop (sqs.h:348-455), multithreaded
    with the shared RNG state, on a bounded sample of the same workload."""
    import numpy as np
    from oracle import oracle as O

    info = job.problem(0)
    prob = O.Problem(info["pairs"].astype(np.uint64), info["species_packed"],
                     np.array(info["bounds"]).ravel(), info["prefactors"], info["pair_weights"],
                     info["target"], seed=7)
    share = cpu_share()
    threads = share["threads"]
    st, _ = O.run([prob], 2000, chunk_size=2000, threads=1, collect=False)
    rate1 = st.finished / max(st.seconds, 1e-9)
    # one thread, about a tenth of the budget
    n1 = max(2000, int(rate1 * budget_s / 10))
    s1, _ = O.run([prob], n1, chunk_size=max(1000, n1 // 8), threads=1, collect=False)
    n = max(threads * 1000, int(rate1 * threads * budget_s))
    chunk = max(1000, n // (threads * 8))
    st, _ = O.run([prob], n, chunk_size=chunk, threads=threads, shared_rng=True, collect=False)
    # per-thread RNG streams (no shared state): the stronger CPU figure, a third of the budget
    n3 = max(threads * 1000, n // 3)
    s3, _ = O.run([prob], n3, chunk_size=max(1000, n3 // (threads * 8)), threads=threads,
                  shared_rng=False, collect=False)
    return {"value": st.finished / st.seconds, "unit": "candidates/s", "cores": threads,
            "kind": "port", "host": dict(share, cpu_model=cpu_model()),
            "sample": f"{st.finished} MC candidates of the same 256-site job, {threads} threads, "
                      f"shared RNG state, {st.seconds:.1f} s on {cpu_model()}",
            "one_thread": {"value": s1.finished / s1.seconds, "cores": 1,
                           "sample": f"{s1.finished} candidates, {s1.seconds:.1f} s"},
            "per_thread_rng": {"value": s3.finished / s3.seconds, "cores": threads,
                               "sample": f"{s3.finished} candidates, {threads} threads with "
                                         f"their own RNG streams, {s3.seconds:.1f} s"}}


class _Hip:
    """the few HIP runtime calls the explicit-scoring measurement needs (device buffers, a
    stream of its own, events), through ctypes"""

    def __init__(self):
        import ctypes as C
        self.C = C
        try:
            self.h = C.CDLL("libamdhip64.so")
        except OSError:
            self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.h.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]

    def check(self, e, what):
        if e != 0:
            raise RuntimeError(f"{what} failed with HIP error {e}")

    def malloc(self, n):
        p = self.C.c_void_p()
        self.check(self.h.hipMalloc(self.C.byref(p), self.C.c_size_t(n)), "hipMalloc")
        return p

    def upload(self, dst, arr):
        self.check(self.h.hipMemcpy(dst, arr.ctypes.data_as(self.C.c_void_p),
                                    self.C.c_size_t(arr.nbytes), 1), "hipMemcpy")

    def download(self, arr, src):
        self.check(self.h.hipMemcpy(arr.ctypes.data_as(self.C.c_void_p), src,
                                    self.C.c_size_t(arr.nbytes), 2), "hipMemcpy")

    def stream(self):
        s = self.C.c_void_p()
        self.check(self.h.hipStreamCreateWithFlags(self.C.byref(s), 1), "hipStreamCreate")
        return s

    def event(self):
        e = self.C.c_void_p()
        self.check(self.h.hipEventCreate(self.C.byref(e)), "hipEventCreate")
        return e


def large_cell_scoring(launches: int, batch: int = 4096) -> dict:
    """explicit scoring of `batch` uniformly random candidates per launch (seed 7) on the FCC
    4x4x4 (256 sites) and 6x6x6 (864 sites) quinary 7-shell cells, by the sparse kernels
    (wavefront per candidate over the LDS-staged pair list; lane per candidate) and the dense
    MFMA kernel; device-resident buffers, events on the launch stream; bond counts of the
    three kernels compared bit for bit"""
    import ctypes as C
    import numpy as np
    import sqsgenerator_amd._lib as L
    import workloads as W
    from sqsgenerator_amd.core import Job, parse_config

    hip = _Hip()
    stream = hip.stream()  # default priority; a NULL stream would mean the job's own stream
    e0, e1 = hip.event(), hip.event()
    out = {}
    for name, make in (("fcc256", W.quinary_256), ("fcc864", W.quinary_864)):
        job = Job(parse_config(make(iterations=batch)))
        info = job.problem(0)
        N, K, S = info["num_sites"], info["num_species"], info["num_shells"]
        rng = np.random.default_rng(7)
        cands = np.repeat(info["species_packed"][None, :], batch, axis=0)
        for b in range(batch):
            for lo, hi in info["bounds"]:
                cands[b, lo:hi] = rng.permutation(cands[b, lo:hi])
        cands = np.ascontiguousarray(cands, dtype=np.uint8)
        d_sp = hip.malloc(cands.nbytes)
        hip.upload(d_sp, cands)
        d_bo = hip.malloc(batch * S * K * K * 4)
        d_ob = hip.malloc(batch * 8)
        res, bonds = {}, {}
        for kname, kernel in (("wave", L.KERNEL_SPARSE_WAVE), ("lane", L.KERNEL_SPARSE_LANE),
                              ("dense", L.KERNEL_DENSE_MFMA)):
            def launch():
                L.check(L.lib().sqsg_score_batch_device(job.handle, 0, d_sp, batch, d_bo, d_ob,
                                                        None, kernel, stream))
            launch()
            hip.check(hip.h.hipStreamSynchronize(stream), "hipStreamSynchronize")
            hip.check(hip.h.hipEventRecord(e0, stream), "hipEventRecord")
            for _ in range(launches):
                launch()
            hip.check(hip.h.hipEventRecord(e1, stream), "hipEventRecord")
            hip.check(hip.h.hipEventSynchronize(e1), "hipEventSynchronize")
            ms = C.c_float()
            hip.check(hip.h.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
            per = ms.value / launches
            res[kname] = {"candidates_per_s": batch / (per * 1e-3), "ms_per_launch": per}
            host = np.empty((batch, S, K, K), dtype=np.int32)
            hip.download(host, d_bo)
            bonds[kname] = host
        res["bonds_identical"] = bool(np.array_equal(bonds["wave"], bonds["dense"]) and
                                      np.array_equal(bonds["lane"], bonds["dense"]))
        res["sites"], res["dense_shape"] = N, job.kernel_shape()
        for p in (d_sp, d_bo, d_ob):
            hip.h.hipFree(p)
        job.close()
        out[name] = res
    for e in (e0, e1):
        hip.h.hipEventDestroy(e)
    hip.h.hipStreamDestroy(stream)
    return out


def autotuned_runs(iterations: int) -> dict:
    """single-job sqsg_run of the benchmark cells at the default launch parameters and after
    sqsg_autotune (variant, dense launch shape, candidate passes per round, launch size):
    candidates per second of the run's wall clock (best of four runs, the default and the
    tuned job alternating), tuning time apart"""
    import sqsgenerator_amd._lib as L
    import workloads as W
    from sqsgenerator_amd.core import Job, parse_config

    cells = [("a_fcc108_ternary", W.ternary_108, iterations),
             ("b_bcc128_quaternary", W.quaternary_bcc_128, iterations),
             ("c_hcp128_binary", W.hcp_128, iterations),
             ("fcc256_quinary", W.quinary_256, iterations),
             ("fcc864_quinary", W.quinary_864, iterations // 8)]

    def run(job):
        L.check(L.lib().sqsg_device_synchronize())
        t0 = time.perf_counter()
        job.run()
        return job.timing()[2] / (time.perf_counter() - t0)

    out = {"iterations": iterations}
    for name, make, its in cells:
        job = Job(parse_config(make(iterations=its)))   # default launch parameters
        tuned = Job(parse_config(make(iterations=its)))
        run(job)  # inputs resident, kernels loaded
        t0 = time.perf_counter()
        tuned.autotune()
        tune_s = time.perf_counter() - t0
        # the two alternate, each first every other time (a short run right after the host
        # work between runs sees a lower clock), best of four each
        rates = {"default": 0.0, "tuned": 0.0}
        for i in range(4):
            for key, j in (("default", job), ("tuned", tuned))[::1 if i % 2 == 0 else -1]:
                rates[key] = max(rates[key], run(j))
        out[name] = {"candidates": its, "default_candidates_per_s": rates["default"],
                     "default_shape": job.kernel_shape(),
                     "tuned_candidates_per_s": rates["tuned"],
                     "tuned_shape": tuned.kernel_shape(), "tuned_launch": tuned.dense_launch(),
                     "tuning_s": tune_s, "speedup": rates["tuned"] / rates["default"]}
        job.close()
        tuned.close()
    return out


def batched_uneven(iterations: int) -> dict:
    """configurations (a)-(d): FCC 108 ternary, BCC 128 quaternary, HCP 128 binary and the
    rocksalt 216 split job, `iterations` MC candidates each, in one batched run (inputs
    resident before the clock starts); each job's batched rate alone beside it"""
    import sqsgenerator_amd._lib as L
    import workloads as W
    from sqsgenerator_amd.core import Job, batch_run, parse_config

    makers = [("a_fcc108_ternary_3shells", W.ternary_108),
              ("b_bcc128_quaternary_4shells", W.quaternary_bcc_128),
              ("c_hcp128_binary_5shells", W.hcp_128),
              ("d_rocksalt216_split_3+3shells", W.rocksalt_split_216)]

    def make(fn):
        job = Job(parse_config(fn(iterations=iterations)))
        for q in range(job.num_problems()):  # upload inputs (and the dense tile stream)
            job.score(job.problem(q)["species_packed"][None, :], q,
                      kernel=L.KERNEL_DENSE_MFMA if job.kernel_shape().startswith("dense")
                      else L.KERNEL_SPARSE_LANE)
        return job

    def timed(jobs):
        L.check(L.lib().sqsg_device_synchronize())
        t0 = time.perf_counter()
        rep = batch_run(jobs)
        L.check(L.lib().sqsg_device_synchronize())
        return time.perf_counter() - t0, rep

    out = {"iterations_per_job": iterations, "jobs": []}
    for name, fn in makers:
        job = make(fn)
        wall, rep = timed([job])
        out["jobs"].append({"job": name, "shape": job.kernel_shape(),
                            "alone_candidates_per_s": rep["candidates"] / wall,
                            "alone_kernel_ms": rep["kernel_ms"], "alone_steps": rep["launches"]})
        job.close()
    # the batched run of all four, once untimed (the engine's groups, graphs and pooled
    # buffers for this mix), then three times over fresh jobs (a run lasts ~15 ms, so the
    # median of three is reported)
    warm = [make(fn) for _, fn in makers]
    timed(warm)
    for j in warm:
        j.close()
    runs = []
    for _ in range(3):
        jobs = [make(fn) for _, fn in makers]
        wall, rep = timed(jobs)
        for j in jobs:
            j.close()
        runs.append((rep["candidates"] / wall, wall, rep))
    runs.sort(key=lambda r: r[0])
    rate, wall, rep = runs[1]
    alone_s = sum(iterations / j["alone_candidates_per_s"] for j in out["jobs"])
    out.update({"together_candidates_per_s": rate, "together_wall_s": wall,
                "together_kernel_ms": rep["kernel_ms"], "together_steps": rep["launches"],
                "together_runs_candidates_per_s": [r[0] for r in runs],
                "sum_of_alone_wall_s": alone_s})
    return out


def serving_jobs(n, seed, max_budget):
    """request mix of the serving stream: FCC/BCC/HCP cells of 32-864 sites, 2-6 species
    with random integer compositions, 2-7 shells (weights 1/s), log-uniform budgets"""
    import numpy as np
    import workloads as W
    from sqsgenerator_amd.core import parse_config

    rng = np.random.default_rng(seed)
    cells = [("fcc", (2, 2, 2)), ("fcc", (2, 2, 4)), ("fcc", (3, 3, 3)), ("bcc", (4, 4, 4)),
             ("hcp", (4, 4, 4)), ("fcc", (4, 4, 4)), ("fcc", (5, 5, 5)), ("fcc", (6, 6, 6))]
    metals = ["Fe", "Co", "Ni", "Cr", "Mn", "Al"]
    out = []
    for _ in range(n):
        lat, cell = cells[rng.integers(len(cells))]
        nsites = {"fcc": 4, "bcc": 2, "hcp": 2}[lat] * int(np.prod(cell))
        K = int(rng.integers(2, 7))
        cuts = np.sort(rng.choice(np.arange(1, nsites), K - 1, replace=False))
        amounts = np.diff(np.concatenate([[0], cuts, [nsites]])).tolist()
        comp = {metals[k]: int(a) for k, a in enumerate(amounts)}
        S = int(rng.integers(2, 8))
        budget = int(10 ** rng.uniform(5, np.log10(max_budget)))
        if lat == "fcc":
            d = W.fcc(4.05, cell, comp, S)
        elif lat == "bcc":
            d = W.bcc(3.165, cell, comp, S)
        else:
            d = W.hcp(2.95, 1.633, cell, comp, S)
        d.update(iterations=budget, chunk_size=min(budget, 100000))
        cfg = parse_config(d)
        while hasattr(cfg, "msg") and S > 2:  # fewer detectable shells than asked
            S -= 1
            d["shell_weights"] = {str(s): 1.0 / s for s in range(1, S + 1)}
            cfg = parse_config(d)
        out.append(cfg)
    return out


def serving(rate, duration, seed, max_budget) -> dict:
    import numpy as np
    from sqsgenerator_amd.core import Job, Server

    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=int(rate * duration * 1.5) + 2))
    arrivals = arrivals[arrivals < duration]
    jobs = [Job(cfg) for cfg in serving_jobs(len(arrivals), seed + 1, max_budget)]
    srv = Server(max_active=64, slice_candidates=1 << 18)
    t0 = time.perf_counter()
    tickets, k = [], 0
    while True:
        now = time.perf_counter() - t0
        while k < len(jobs) and arrivals[k] <= now:
            tickets.append(srv.submit(jobs[k]))
            k += 1
        active, queued = srv.step()
        if k == len(jobs) and active == 0 and queued == 0:
            break
        if active == 0 and queued == 0:
            time.sleep(max(0.0, min(0.005, arrivals[k] - (time.perf_counter() - t0))))
    wall = time.perf_counter() - t0
    lat, cands = [], 0
    for t, j in zip(tickets, jobs):
        st = srv.poll(t)
        lat.append(st["t_done"] - st["t_submit"])
        cands += j.timing()[2]
    srv.close()
    for j in jobs:
        j.close()
    lat = np.array(lat)
    window = float(arrivals[-1]) if len(arrivals) else duration
    return {"arrival_rate_jobs_per_s": rate, "jobs": len(jobs), "duration_s": wall,
            "p50_latency_s": float(np.percentile(lat, 50)),
            "p99_latency_s": float(np.percentile(lat, 99)),
            "aggregate_candidates_per_s": cands / wall,
            # the candidates the requests bring per second of the arrival window: the
            # aggregate cannot exceed it while the device keeps up
            "offered_candidates_per_s": cands / max(window, 1e-9),
            "mix": "FCC/BCC/HCP 32-864 sites, 2-6 species, 2-7 shells, MC budgets log-uniform "
                   f"1e5-{max_budget:.0e}, Poisson arrivals (seed {seed}); requests parsed "
                   "before the clock starts, latency = completion - submission"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--candidates-per-step", type=int, default=1 << 23,
                    help="candidates of each step (a launch of the plain kernel, or one "
                         "persistent step over the batched jobs); a persistent step ends "
                         "with its slowest wavefront's last 60-candidate segment, so longer "
                         "steps waste less of the launch")
    ap.add_argument("--batch-jobs", type=int, default=4)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--serve-rate", type=float, default=30.0)
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-max-budget", type=float, default=1e7)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--large-cells", action="store_true",
                    help="only the large-cell sparse-vs-dense scoring (4096 candidates per launch)")
    ap.add_argument("--large-cell-launches", type=int, default=10)
    ap.add_argument("--no-large-cells", action="store_true")
    ap.add_argument("--uneven-iterations", type=int, default=10_000_000,
                    help="MC candidates per job of the batched uneven leg (configs (a)-(d)); "
                         "0: skip it")
    ap.add_argument("--autotune-iterations", type=int, default=16_000_000,
                    help="MC candidates per single-job run of the autotuner leg (1/8 of it at "
                         "864 sites); 0: skip it")
    ap.add_argument("--no-prune", action="store_true",
                    help="skip the secondary run with early exit (sqsg_job_set_pruning)")
    ap.add_argument("--serve-sweep", default="10,30,100,300,1000",
                    help="arrival rates (jobs/s) of the serving sweep; empty: none")
    ap.add_argument("--serve-sweep-jobs", type=int, default=240,
                    help="requests per sweep rate at most (the window shrinks with the rate)")
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
        print("bench.py measures one GPU only (candidates are independent, so N GPUs would "
              "each run disjoint RNG streams / rank ranges and merge their tie sets on the "
              "host; that merge is not built)", file=sys.stderr)
        return 2

    import sqsgenerator_amd._lib as L
    import workloads as W
    from sqsgenerator_amd.core import Job, batch_run, parse_config

    if args.large_cells:
        print(json.dumps({"large_cell_scoring": large_cell_scoring(args.large_cell_launches)}),
              flush=True)
        return 0
    per_step = args.candidates_per_step
    nb = args.batch_jobs

    def make(iterations, seed=7, kernel=L.KERNEL_AUTO, launch=per_step):
        job = Job(parse_config(W.quinary_256(iterations=iterations, seed=seed)))
        job.set_kernel(kernel)
        job.set_launch_size(launch)
        return job

    # plain single-job kernel
    plain = None
    if args.plain_steps:
        pj = make(per_step, kernel=L.KERNEL_SPARSE_WAVE)
        pj.run()
        pj.close()
        pj = make(args.plain_steps * per_step, kernel=L.KERNEL_SPARSE_WAVE)
        pj.run()
        pms, plaunch, pn = pj.timing()
        plain = {"kernel_candidates_per_s": pn / (pms / 1e3), "avg_launch_ms": pms / plaunch}
        pj.close()

    # batched persistent form: nb jobs advanced by one persistent launch per step
    def batch_jobs(steps, prune=False):
        jobs = [make(steps * per_step // nb, seed=7 + k, launch=per_step // nb)
                for k in range(nb)]
        for j in jobs:
            j.set_kernel(L.KERNEL_AUTO)
            j.set_pruning(prune)
        for j in jobs:
            # inputs resident in HBM: pair lists, parameter cubes and the dense tile stream
            # (explicit scoring through the dense kernel builds and uploads it)
            j.score(j.problem(0)["species_packed"][None, :], kernel=L.KERNEL_DENSE_MFMA)
        return jobs

    # the timed jobs are built and their inputs uploaded first, so that the warm-up steps
    # run right before the timed ones (the device is at its steady clock when timing starts)
    jobs = batch_jobs(args.steps)
    info = jobs[0].problem(0)
    shape = jobs[0].kernel_shape()  # e.g. dense<8,2,2,3> columns 4
    kernel_name = ("batch_dense_kernel" + shape[len("dense"):] if shape.startswith("dense")
                   else shape)
    warm = batch_jobs(args.warmup) if args.warmup else []
    if warm:
        batch_run(warm)
    L.check(L.lib().sqsg_device_synchronize())
    t0 = time.perf_counter()
    rep = batch_run(jobs)
    L.check(L.lib().sqsg_device_synchronize())
    wall = time.perf_counter() - t0
    scored = rep["candidates"]
    assert scored == args.steps * per_step, (scored, args.steps * per_step)
    for j in jobs + warm:
        j.close()
    value = scored / wall
    pruned = None
    if not args.no_prune:
        # the same jobs with early exit of passes that cannot reach the kept levels; the
        # result packs are identical (tests), so this is time to solution, not the headline
        pj = batch_jobs(args.steps, prune=True)
        pwarm = batch_jobs(args.warmup, prune=True) if args.warmup else []
        if pwarm:
            batch_run(pwarm)
        L.check(L.lib().sqsg_device_synchronize())
        t1 = time.perf_counter()
        prep = batch_run(pj)
        L.check(L.lib().sqsg_device_synchronize())
        pwall = time.perf_counter() - t1
        for j in pj + pwarm:
            j.close()
        pruned = {"candidates_per_s": prep["candidates"] / pwall,
                  "kernel_candidates_per_s": prep["candidates"] / (prep["kernel_ms"] / 1e3),
                  "note": "same jobs with sqsg_job_set_pruning (a pass stops once none of its "
                          "candidates can reach the kept levels; identical result packs); "
                          "not the headline"}
    avg_launch_s = rep["kernel_ms"] / 1e3 / rep["launches"]
    P, N = info["num_pairs"], info["num_sites"]
    K, S = info["num_species"], info["num_shells"]
    per_launch = scored // rep["launches"]
    ops_per_cand = dense_ops_per_candidate(N, K, S)
    achieved = ops_per_cand * per_launch / avg_launch_s / 1e12
    kernel_rate = scored / (rep["kernel_ms"] / 1e3)
    out = {
        "metric": METRIC, "value": value, "unit": "candidates/s", "n_gpus": 1,
        "steps": rep["launches"], "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / rep["launches"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i8 MFMA->i32 counts, f64 objective",
        "data": "synthetic (random MC candidates of the generated cell)",
        "config": {"workload": "fcc4x4x4-256site-quinary-7shell-mc", "sites": N,
                   "species": info["num_species"], "shells": info["num_shells"], "pairs": P,
                   "form": f"batched persistent, {nb} jobs (seeds 7..{6 + nb}), graph-captured",
                   "candidates_per_step": per_step,
                   "kernel_candidates_per_s": kernel_rate,
                   "plain_kernel": plain,
                   "pruned": pruned,
                   "speedup_vs_plain_kernel":
                       kernel_rate / plain["kernel_candidates_per_s"] if plain else None},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": INT8_PEAK_TOPS,
                     "unit": "TFLOP/s", "frac": achieved / INT8_PEAK_TOPS,
                     "traffic": measured_traffic(per_launch),
                     "kernel": kernel_name, "avg_launch_ms": avg_launch_s * 1e3,
                     "algorithmic_ops_per_candidate": ops_per_cand,
                     "executed_ops_per_candidate": (executed_ops_per_candidate(shape, S)
                                                    if " columns " in shape else None),
                     "candidates_per_launch": per_launch,
                     "note": "int8 ops (2 per multiply-accumulate) of the dense one-hot "
                             "contraction S(2N^2K + 2NK^2) per candidate; the kernel executes "
                             "fewer (upper tiles, the last species without a column); launch "
                             "time from HIP events around the graph launch on the engine's "
                             "stream"},
    }
    if not args.no_large_cells:
        out["large_cell_scoring"] = large_cell_scoring(args.large_cell_launches)
    if args.uneven_iterations:
        out["batched_uneven"] = batched_uneven(args.uneven_iterations)
    if args.autotune_iterations:
        out["autotune"] = autotuned_runs(args.autotune_iterations)
    if not args.no_serve:
        out["serving"] = serving(args.serve_rate, args.serve_seconds, 20261203,
                                 args.serve_max_budget)
        sweep = []
        for r in (float(x) for x in args.serve_sweep.split(",") if x.strip()):
            window = min(args.serve_seconds, args.serve_sweep_jobs / r)
            res = serving(r, window, 20261203, args.serve_max_budget)
            res.pop("mix")
            sweep.append(res)
        if sweep:
            out["serving_sweep"] = sweep
    if not args.no_cpu:
        cj = make(per_step)
        out["cpu_baseline"] = cpu_baseline(cj, args.cpu_budget)
        cj.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
