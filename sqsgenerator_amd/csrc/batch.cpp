// Batched persistent execution and continuous-batching service.
//
// BatchEngine keeps a set of resident jobs (their inputs, thresholds and record sinks stay
// in device memory) and advances all of them with one persistent launch per step: each
// step gives every active job its next slice of candidates as one work item, and the
// persistent wavefronts claim the items one segment at a time through per-item counters.
// Jobs are grouped by kernel variant and its specialisation (largest species count for the
// lane kernel, tile bound for the dense MFMA kernel); each group has fixed device tables
// and a captured graph {counter reset kernel, persistent kernel} that is replayed every
// step.  Steps are pipelined two deep: each job's record sink and counters are
// double-buffered by step parity, so while step k runs the host reads back step k - 1,
// merges its records, tightens each job's threshold and admits or retires jobs (which is
// what the serving loop builds on); the device never waits for the host between steps.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <condition_variable>
#include <exception>
#include <thread>
#include <atomic>

#include "job.hpp"
#include "kernels.hpp"

namespace sqsg {

namespace {

// held around graph captures and around device uploads made on other host threads
std::mutex &capture_mutex() {
  static std::mutex m;
  return m;
}

constexpr uint32_t kGroupJobs = 256;          // resident jobs per group
constexpr uint32_t kGroupItems = 2 * kGroupJobs;  // a job's re-issued range + its next slice

struct Group;

using Range = std::pair<uint64_t, uint64_t>;

// Device buffers of a resident job, pooled by the engine across jobs: record sinks and
// counters double-buffered per step parity (a step writes one set while the host reads the
// other), and the stages in pinned host memory that the stage kernel fills at the end of
// each step (counts and the first records).
struct SlotBuffers {
  RecordSink sink[2];
  DevBuf count[2], scored[2];
  uint8_t *stage[2] = {nullptr, nullptr};      // host addresses
  uint8_t *stage_dev[2] = {nullptr, nullptr};  // the same memory as the device sees it
  size_t stage_cap = 0;
  SlotBuffers() {
    for (int q = 0; q < 2; ++q) {
      count[q].alloc(sizeof(unsigned int));
      scored[q].alloc(sizeof(unsigned long long));
    }
  }
  void ensure_stage(size_t bytes) {
    if (bytes <= stage_cap) return;
    release_stage();
    for (int q = 0; q < 2; ++q) {
      void *h = nullptr, *d = nullptr;
      SQSG_HIP(hipHostMalloc(&h, bytes, hipHostMallocMapped | hipHostMallocCoherent));
      stage[q] = static_cast<uint8_t *>(h);
      SQSG_HIP(hipHostGetDevicePointer(&d, h, 0));
      stage_dev[q] = static_cast<uint8_t *>(d);
    }
    stage_cap = bytes;
  }
  void release_stage() {
    for (int q = 0; q < 2; ++q) {
      if (stage[q]) (void)hipHostFree(stage[q]);
      stage[q] = stage_dev[q] = nullptr;
    }
    stage_cap = 0;
  }
  ~SlotBuffers() { release_stage(); }
};

struct Slot {
  sqsg_job *job = nullptr;
  Group *group = nullptr;
  RunParams p{};
  std::unique_ptr<SlotBuffers> buf;
  RecordSink *sink = nullptr;   // buf->sink
  DevBuf *count = nullptr, *scored = nullptr;
  uint64_t next = 0;          // first candidate not issued yet
  uint64_t slice = 0;         // candidates per step
  uint32_t ramp = 0;          // steps issued so far
  bool ramp_up = false;       // early steps are shorter (records not bounded by the filter)
  uint64_t chunk = 0;         // candidates a wavefront claims at once (one segment)
  double cost = 0;            // relative device work per candidate
  std::vector<Range> issued[2];  // ranges of the step of each parity
  std::deque<Range> retry;    // ranges to issue again (their record sink overflowed)
  size_t footprint = 0;       // device bytes held
  bool done = false;
};

struct Group {
  int variant = 0;
  uint32_t key = 0;    // the group's identity within its variant
  uint32_t shape = 0;  // launch template: lane: largest species count; dense: tile bound
  uint32_t alt = 0;    // dense: the alternative shape of the tile bound (0: default)
  DevBuf d_jobs, d_items, d_counts, d_claims;
  hipGraphExec_t exec = nullptr;
  uint32_t blocks = 0;     // of the current step's launch
  BatchParams bp{};
  // pinned host sources of each parity's uploads (asynchronous copies, and the buffer of
  // a step stays untouched until that step was processed)
  RunParams *h_table[2] = {nullptr, nullptr};
  WorkItem *h_items[2] = {nullptr, nullptr};
  uint32_t *h_counts[2] = {nullptr, nullptr};
  uint32_t graph_wave_lds = 0, graph_blocks = 0;
  double work = 0;  // estimated cost of the current step (candidates x pair work)
  bool split = false;  // dense: the kernel form of split-mode jobs
  std::vector<Slot *> members;
  Group() {
    for (int q = 0; q < 2; ++q) {
      void *t = nullptr, *i = nullptr, *c = nullptr;
      SQSG_HIP(hipHostMalloc(&t, kGroupJobs * sizeof(RunParams), hipHostMallocDefault));
      h_table[q] = static_cast<RunParams *>(t);
      SQSG_HIP(hipHostMalloc(&i, kGroupItems * sizeof(WorkItem), hipHostMallocDefault));
      h_items[q] = static_cast<WorkItem *>(i);
      SQSG_HIP(hipHostMalloc(&c, 2 * sizeof(uint32_t), hipHostMallocDefault));
      h_counts[q] = static_cast<uint32_t *>(c);
    }
  }
  ~Group() {
    if (exec) (void)hipGraphExecDestroy(exec);
    for (int q = 0; q < 2; ++q) {
      if (h_table[q]) (void)hipHostFree(h_table[q]);
      if (h_items[q]) (void)hipHostFree(h_items[q]);
      if (h_counts[q]) (void)hipHostFree(h_counts[q]);
    }
  }
};

size_t job_footprint(const sqsg_job *job) {
  size_t b = 0;
  for (const auto &P : job->problems)
    b += P.pi.size() * 4 + P.lane_stream.size() * 4 + P.pref.size() * 24 + 2 * P.N + 1024 +
         size_t(P.dtiles_n) * 1024 + 32 * P.dT;
  return b;
}

}  // namespace

class BatchEngine {
 public:
  explicit BatchEngine(bool use_graph) : use_graph_(use_graph) {
    SQSG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    SQSG_HIP(hipStreamCreateWithFlags(&stream2_, hipStreamNonBlocking));
    SQSG_HIP(hipStreamCreateWithFlags(&copy_, hipStreamNonBlocking));
    for (int q = 0; q < 2; ++q) {
      SQSG_HIP(hipEventCreate(&ev_start_[q]));
      SQSG_HIP(hipEventCreate(&ev_done_[q]));
      SQSG_HIP(hipEventCreateWithFlags(&ev_join_[q], hipEventDisableTiming));
    }
  }
  ~BatchEngine() {
    if (stream_) (void)hipStreamSynchronize(stream_);
    groups_.clear();
    slots_.clear();
    retired_[0].clear();
    retired_[1].clear();
    pool_.clear();
    for (int q = 0; q < 2; ++q) {
      if (ev_start_[q]) (void)hipEventDestroy(ev_start_[q]);
      if (ev_done_[q]) (void)hipEventDestroy(ev_done_[q]);
      if (ev_join_[q]) (void)hipEventDestroy(ev_join_[q]);
    }
    if (copy_) (void)hipStreamDestroy(copy_);
    if (stream2_) (void)hipStreamDestroy(stream2_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // Makes a job resident; its run state is reset exactly as sqsg_run would.
  size_t add(sqsg_job *job) {
    begin_run(job);
    const int variant = resolve_kernel(job, job->kernel == SQSG_KERNEL_SPARSE_WAVE
                                                ? SQSG_KERNEL_AUTO : job->kernel);
    if (variant == SQSG_KERNEL_SPARSE_WAVE)
      throw std::invalid_argument("batched execution needs a job the lane or dense kernel "
                                  "handles");
    const bool dense = variant == SQSG_KERNEL_DENSE_MFMA || variant == SQSG_KERNEL_DENSE_PAIRED;
    if (dense) ensure_dense(job);
    auto s = std::make_unique<Slot>();
    s->job = job;
    uint32_t blocks = 0;
    s->p = make_run_params(job, variant, 1, &blocks);
    s->slice = job->launch_size ? job->launch_size : default_step_size(job);
    uint64_t work = 0;
    for (const auto &P : job->problems) work += P.pi.size() + 64;
    const uint64_t gen = s->p.gen_group;
    if (s->p.systematic) {
      s->p.rounds = static_cast<uint32_t>(std::clamp<uint64_t>(
          (uint64_t(1) << 21) / (gen * std::max<uint64_t>(work, 1)), 1, 256));
    } else {
      s->p.rounds = 1;
    }
    // a paired block claims one batch for its four pairs at a time
    s->chunk = gen * s->p.rounds * (variant == SQSG_KERNEL_DENSE_PAIRED ? kDensePairs : 1);
    if (dense) {  // every grid position of the tile bound costs its MFMAs in every shell
      const double TT = dense_tile_bound_of(job);
      double shells = 0;
      for (const auto &P : job->problems) shells += P.S;
      s->cost = shells * TT * (TT + 1) / 2 / s->p.dense_cw;
    } else {
      s->cost = static_cast<double>(work) / 64;
    }
    s->slice = std::max(s->slice, s->chunk);
    s->p.rec_stride = job->stride;
    if (!pool_.empty()) {
      s->buf = std::move(pool_.back());
      pool_.pop_back();
    } else {
      s->buf = std::make_unique<SlotBuffers>();
    }
    s->sink = s->buf->sink;
    s->count = s->buf->count;
    s->scored = s->buf->scored;
    const uint32_t cap =
        first_sink_records(job, uint64_t(device_cu_count()) * 8, s->slice);
    s->ramp_up = !records_bounded(job);
    for (int q = 0; q < 2; ++q) s->sink[q].ensure(cap, job->stride, s->p.nprob);
    // the stage holds a probe's records (the first step of a served job) as well
    s->p.stage_cap = std::min<uint32_t>(cap, 4096);
    s->buf->ensure_stage(stage_bytes(s->p.stage_cap, s->p.nprob, job->stride));
    s->p.thr_key = search_key(job);  // travels in the job table of every step
    s->footprint = footprint(*s);
    // groups share a launch shape: the tile bound (dense), also K (paired: one LDS
    // layout), the largest species count (lane); split jobs have dense kernels of their own
    const bool split = job->problems.size() > 1;
    // (and the tuned alternative shape)
    const uint32_t alt = variant == SQSG_KERNEL_DENSE_MFMA ? search_alt(job) : 0u;
    const uint32_t key = !dense ? job->kmax
                         : variant == SQSG_KERNEL_DENSE_PAIRED
                             ? dense_tile_bound_of(job) | (job->problems[0].K << 8)
                             : dense_tile_bound_of(job) | (split ? 1u << 16 : 0u) | (alt << 20);
    Group &g = group(variant, key, dense ? dense_tile_bound_of(job) : job->kmax);
    g.split = dense && split;
    g.alt = alt;
    if (g.members.size() >= kGroupJobs) throw std::out_of_range("too many resident jobs");
    g.members.push_back(s.get());
    s->group = &g;
    const size_t fp = s->footprint;
    slots_.push_back(std::move(s));
    return fp;
  }

  bool idle() const { return slots_.empty() && !in_flight_[0] && !in_flight_[1]; }
  // releases retired slots (only when idle: no issued table names them any more)
  void release_retired() {
    if (!idle()) return;
    recycle(retired_[0]);
    recycle(retired_[1]);
  }
  size_t resident() const { return slots_.size(); }

  // Issues the next persistent step (when any resident job has work left), then processes
  // the previous step while the new one runs: counters, record sinks, thresholds,
  // retirement.  Finished jobs are finalised and returned.
  std::vector<sqsg_job *> step() {
    const int p = parity_;
    parity_ ^= 1;
    recycle(retired_[p]);  // the last table that listed them was parity p's previous step
    bool work = false;
    for (auto &s : slots_)
      work |= !s->retry.empty() || (s->next < s->job->iterations && !s->job->cancel.load());
    if (work) {
      const auto t0 = std::chrono::steady_clock::now();
      std::vector<Group *> live;
      for (auto &[k, g] : groups_)
        if (!g->members.empty()) live.push_back(prepare_group(*g, p));
      // the groups' persistent kernels run on two streams, so that one group's kernel fills
      // the tail of another's; heaviest group first, each to the stream with less work
      std::sort(live.begin(), live.end(),
                [](const Group *a, const Group *b) { return a->work > b->work; });
      double load[2] = {0, 0};
      SQSG_HIP(hipEventRecord(ev_start_[p], stream_));
      SQSG_HIP(hipStreamWaitEvent(stream2_, ev_start_[p], 0));
      for (Group *g : live) {
        const int k = load[1] < load[0] ? 1 : 0;
        load[k] += g->work;
        launch_group(*g, k ? stream2_ : stream_);
      }
      SQSG_HIP(hipEventRecord(ev_join_[p], stream2_));
      SQSG_HIP(hipStreamWaitEvent(stream_, ev_join_[p], 0));
      SQSG_HIP(hipEventRecord(ev_done_[p], stream_));
      in_flight_[p] = true;
      host_ns_ += std::chrono::duration_cast<std::chrono::nanoseconds>(
                      std::chrono::steady_clock::now() - t0).count();
    }
    // the previous call processed the step before that, so only parity p ^ 1 can be pending
    std::vector<sqsg_job *> finished;
    if (in_flight_[p ^ 1]) process(p ^ 1, finished);
    return finished;
  }

  double kernel_ms = 0;
  uint64_t candidates = 0, launches = 0;
  uint32_t last_items = 0;
  bool probe_new_ = false;  // new jobs start with a probe range (serving)

 private:
  // returns the buffers of retired slots to the pool
  void recycle(std::vector<std::unique_ptr<Slot>> &slots) {
    for (auto &s : slots) pool_.push_back(std::move(s->buf));
    slots.clear();
  }

  size_t footprint(const Slot &s) const {
    return job_footprint(s.job) + (s.sink[0].cap + s.sink[1].cap) * (24 + s.job->stride);
  }

  // Reads back the step of parity q (issued before the one now running) and applies it.
  void process(int q, std::vector<sqsg_job *> &finished) {
    const auto t0 = std::chrono::steady_clock::now();
    SQSG_HIP(hipEventSynchronize(ev_done_[q]));
    in_flight_[q] = false;
    float ms = 0;
    SQSG_HIP(hipEventElapsedTime(&ms, ev_start_[q], ev_done_[q]));
    kernel_ms += ms;
    ++launches;
    std::atomic_thread_fence(std::memory_order_acquire);  // the stages were written
    // records are copied to the host here, then merged (jobs in parallel: each merge
    // touches only its own job's collection)
    struct MergeTask {
      Slot *slot;
      std::shared_ptr<const RecordBlock> blk;
      std::vector<Range> ranges;
    };
    std::vector<MergeTask> tasks;
    size_t total_records = 0;
    for (auto &s : slots_) {
      sqsg_job *job = s->job;
      if (!s->issued[q].empty()) {
        uint64_t expect = 0;
        for (const Range &r : s->issued[q]) expect += r.second - r.first;
        const uint8_t *st = s->buf->stage[q];
        unsigned int count = 0;
        unsigned long long scored = 0;
        std::memcpy(&count, st, sizeof(count));
        std::memcpy(&scored, st + 8, sizeof(scored));
        job->launches++;
        job->kernel_ms += ms;  // shared launch: every resident job saw the whole step
        if (count > s->sink[q].cap) {
          // the sink kept only part of the records: widen it and issue the ranges again
          s->sink[q].ensure(count + count / 4, job->stride, s->p.nprob);
          s->footprint = footprint(*s);
          for (const Range &r : s->issued[q]) s->retry.push_back(r);
        } else {
          if (scored != expect)
            throw DeviceError("persistent launch scored an unexpected candidate count");
          job->scored += expect;
          candidates += expect;
          // many records (a job's first steps) come from the device sink, not the stage
          tasks.push_back({s.get(),
                           count <= s->p.stage_cap
                               ? stage_block(job, st, s->p.stage_cap, count, s->p.nprob)
                               : s->sink[q].fetch(job, count, s->p.nprob, copy_),
                           s->issued[q]});
          total_records += count;
          job->stats.finished = job->scored;
        }
        s->issued[q].clear();
      }
      const bool exhausted = s->next >= job->iterations || job->cancel.load();
      if (exhausted && s->retry.empty() && s->issued[q ^ 1].empty()) s->done = true;
    }
    auto merge = [&tasks](size_t k) { merge_block(tasks[k].slot->job, tasks[k].blk, tasks[k].ranges); };
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (total_records < 16384 || tasks.size() < 2 || hw == 1) {
      for (size_t k = 0; k < tasks.size(); ++k) merge(k);
    } else {
      std::atomic<size_t> next{0};
      std::vector<std::thread> pool;
      std::exception_ptr err;
      std::mutex err_mutex;
      for (unsigned t = 0; t < std::min<size_t>(hw, tasks.size()); ++t)
        pool.emplace_back([&] {
          for (size_t k; (k = next.fetch_add(1)) < tasks.size();) {
            try {
              merge(k);
            } catch (...) {
              std::lock_guard<std::mutex> lock(err_mutex);
              err = std::current_exception();
            }
          }
        });
      for (auto &t : pool) t.join();
      if (err) std::rethrow_exception(err);
    }
    // the host's threshold after the candidates merged so far goes into the job table of
    // the next step (the step running now keeps the older, looser one)
    for (const MergeTask &t : tasks) t.slot->p.thr_key = search_key(t.slot->job);
    const auto t1 = std::chrono::steady_clock::now();
    const int64_t ns =
        std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count() + host_ns_;
    host_ns_ = 0;
    for (auto it = slots_.begin(); it != slots_.end();) {
      (*it)->job->stats.timing_loop_ns += ns;
      if (!(*it)->done) {
        ++it;
        continue;
      }
      sqsg_job *job = (*it)->job;
      // SRO cubes are re-scored when first read: device allocations here would wait for the
      // step that is running
      finalize_results(job, false);
      job->stats.timing_total_ns = job->stats.timing_loop_ns;
      Group &g = *(*it)->group;
      g.members.erase(std::find(g.members.begin(), g.members.end(), it->get()));
      finished.push_back(job);
      // its counters stay allocated while an issued table may still name them
      retired_[q ^ 1].push_back(std::move(*it));
      it = slots_.erase(it);
    }
  }

  Group &group(int variant, uint32_t key, uint32_t shape) {
    auto &g = groups_[(static_cast<uint64_t>(variant) << 32) | key];
    if (!g) {
      g = std::make_unique<Group>();
      g->variant = variant;
      g->key = key;
      g->shape = shape;
      g->d_jobs.alloc(kGroupJobs * sizeof(RunParams));
      g->d_items.alloc(kGroupItems * sizeof(WorkItem));
      g->d_counts.alloc(2 * sizeof(uint32_t));
      g->d_claims.alloc(kGroupItems * sizeof(unsigned int));
    }
    return *g;
  }

  // Builds and uploads the group's job table and work items for a step of parity q (the
  // uploads are ordered after the step running now, which still reads the device tables)
  Group *prepare_group(Group &g, int q) {
    RunParams *table = g.h_table[q];
    WorkItem *items = g.h_items[q];
    uint32_t n = 0;
    g.work = 0;
    uint32_t wave_lds = 0;
    for (uint32_t k = 0; k < g.members.size(); ++k) {
      Slot *s = g.members[k];
      s->sink[q].bind(s->p);
      s->p.rec_count = s->count[q].as<unsigned int>();
      s->p.scored = s->scored[q].as<unsigned long long>();
      s->p.stage = s->buf->stage_dev[q];
      table[k] = s->p;
      wave_lds = std::max(wave_lds, s->p.wave_lds);
      s->issued[q].clear();
      if (!s->retry.empty()) {
        s->issued[q].push_back(s->retry.front());
        s->retry.pop_front();
      }
      // a job admitted to a running engine first scores a probe of 4096 candidates and
      // takes its next range once the probe is merged, so that its threshold is set: with
      // the threshold still unset, every wavefront records its first candidates of the job
      const bool waiting = probe_new_ && s->ramp == 1 && s->job->coll.merged < s->next;
      if (s->next < s->job->iterations && !s->job->cancel.load() && !waiting) {
        // without the wavefront filter the device records every candidate at or below the
        // threshold as it stood before the step: such steps grow from 4096 candidates by 4x
        // up to the slice
        const uint64_t ramped = std::max(s->chunk, uint64_t(4096) << std::min(2 * s->ramp, 40u));
        const uint64_t size =
            std::min(s->slice, s->ramp_up || (probe_new_ && s->ramp == 0) ? ramped : s->slice);
        ++s->ramp;
        const uint64_t end = std::min(s->job->iterations, s->next + size);
        s->issued[q].push_back(Range{s->next, end});
        s->next = end;
      }
      for (const Range &r : s->issued[q]) {
        items[n++] = WorkItem{k, static_cast<uint32_t>(s->chunk), r.first, r.second};
        g.work += double(r.second - r.first) * s->cost;
      }
    }
    last_items = n;
    SQSG_HIP(hipMemcpyAsync(g.d_jobs.as<void>(), table, g.members.size() * sizeof(RunParams),
                            hipMemcpyHostToDevice, stream_));
    if (n)
      SQSG_HIP(hipMemcpyAsync(g.d_items.as<void>(), items, n * sizeof(WorkItem),
                              hipMemcpyHostToDevice, stream_));
    g.h_counts[q][0] = n;
    g.h_counts[q][1] = static_cast<uint32_t>(g.members.size());
    SQSG_HIP(hipMemcpyAsync(g.d_counts.as<void>(), g.h_counts[q], 2 * sizeof(uint32_t),
                            hipMemcpyHostToDevice, stream_));
    // a captured graph keeps its LDS size while members come and go, so the size only grows
    // (each change is a new capture)
    if (use_graph_ && g.exec) wave_lds = std::max(wave_lds, g.graph_wave_lds);
    const size_t block_lds = wave_lds;
    uint32_t per_cu = static_cast<uint32_t>((160 * 1024) / std::max<size_t>(block_lds, 1));
    per_cu = std::max(1u, std::min(per_cu, 16u));
    if (g.variant == SQSG_KERNEL_DENSE_MFMA)
      per_cu = g.alt ? dense_alt_waves_per_cu(g.shape, g.alt, block_lds, true)
                     : dense_waves_per_cu(g.shape, block_lds, true, g.split);
    else if (g.variant == SQSG_KERNEL_DENSE_PAIRED)
      per_cu = dense_paired_blocks_per_cu(g.shape, block_lds, true);
    const uint32_t blocks = static_cast<uint32_t>(device_cu_count()) * per_cu;
    BatchParams &bp = g.bp;
    bp = BatchParams{};
    bp.jobs = g.d_jobs.as<const RunParams>();
    bp.items = g.d_items.as<const WorkItem>();
    bp.num_items = g.d_counts.as<const uint32_t>();
    bp.num_jobs = g.d_counts.as<const uint32_t>() + 1;
    bp.wave_lds = wave_lds;
    bp.claims = g.d_claims.as<unsigned int>();
    bp.max_items = kGroupItems;
    bp.stagger_from = blocks / 2;
    g.blocks = blocks;
    SQSG_HIP(g.variant == SQSG_KERNEL_DENSE_MFMA
                 ? (g.alt ? prepare_batch_dense_alt(g.shape, g.alt, block_lds)
                          : prepare_batch_dense(g.shape, block_lds, g.split))
             : g.variant == SQSG_KERNEL_DENSE_PAIRED ? prepare_batch_dense_paired(g.shape, block_lds)
                                                     : prepare_batch_lane(g.shape, block_lds));
    if (use_graph_ && (!g.exec || g.graph_wave_lds != wave_lds || g.graph_blocks != blocks)) {
      // the graph covers every table row (num_jobs is a kernel argument of the reset)
      if (g.exec) SQSG_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
      hipGraph_t graph = nullptr;
      // no synchronous device call of another thread (a server's job preparation) may run
      // while the stream is being captured
      std::lock_guard<std::mutex> capture_lock(capture_mutex());
      SQSG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      const hipError_t e1 = launch_batch_reset(bp, kGroupJobs, stream_);
      const hipError_t e2 = launch_kernel(g, stream_);
      const hipError_t e3s = launch_batch_stage(bp, kGroupJobs, stream_);
      SQSG_HIP(hipStreamEndCapture(stream_, &graph));
      SQSG_HIP(e1);
      SQSG_HIP(e2);
      SQSG_HIP(e3s);
      const hipError_t e3 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      SQSG_HIP(e3);
      g.graph_wave_lds = wave_lds;
      g.graph_blocks = blocks;
    }
    return &g;
  }

  hipError_t launch_kernel(const Group &g, hipStream_t st) {
    switch (g.variant) {
      case SQSG_KERNEL_DENSE_MFMA:
        return g.alt ? launch_batch_dense_alt(g.bp, g.shape, g.alt, g.blocks, st)
                     : launch_batch_dense(g.bp, g.shape, g.blocks, st, g.split);
      case SQSG_KERNEL_DENSE_PAIRED:
        return launch_batch_dense_paired(g.bp, g.shape, g.blocks, st);
      default: return launch_batch_lane(g.bp, g.shape, g.blocks, 1, st);
    }
  }

  // the step's launches only (tables, attributes and graphs are prepared beforehand)
  void launch_group(Group &g, hipStream_t st) {
    if (!use_graph_) {
      SQSG_HIP(launch_batch_reset(g.bp, kGroupJobs, st));
      SQSG_HIP(launch_kernel(g, st));
      SQSG_HIP(launch_batch_stage(g.bp, kGroupJobs, st));
      return;
    }
    SQSG_HIP(hipGraphLaunch(g.exec, st));
  }

  bool use_graph_;
  hipStream_t stream_ = nullptr;   // steps (table uploads, threshold updates, half the groups)
  hipStream_t stream2_ = nullptr;  // the other half of a step's groups
  hipStream_t copy_ = nullptr;     // read-back of finished steps
  hipEvent_t ev_start_[2] = {nullptr, nullptr}, ev_done_[2] = {nullptr, nullptr};
  hipEvent_t ev_join_[2] = {nullptr, nullptr};
  bool in_flight_[2] = {false, false};
  int parity_ = 0;
  int64_t host_ns_ = 0;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::vector<std::unique_ptr<Slot>> retired_[2];
  std::vector<std::unique_ptr<SlotBuffers>> pool_;
  std::map<uint64_t, std::unique_ptr<Group>> groups_;
};

// One engine per graph mode for the whole process: its groups, device tables and captured
// graphs are reused by later calls.  It is never destroyed (the HIP runtime may already be
// torn down when static destructors run).
void batch_run(sqsg_job **jobs, uint32_t n, int use_graph, sqsg_batch_report *report) {
  static std::mutex mutex;
  static BatchEngine *engines[2] = {nullptr, nullptr};
  std::lock_guard<std::mutex> lock(mutex);
  BatchEngine *&eng = engines[use_graph ? 1 : 0];
  if (!eng) eng = new BatchEngine(use_graph != 0);
  const double ms0 = eng->kernel_ms;
  const uint64_t c0 = eng->candidates, l0 = eng->launches;
  try {
    for (uint32_t k = 0; k < n; ++k) eng->add(jobs[k]);
  } catch (...) {
    while (!eng->idle()) eng->step();  // the jobs added so far finish normally
    eng->release_retired();
    throw;
  }
  uint32_t max_items = 0;
  while (!eng->idle()) {
    eng->step();
    max_items = std::max(max_items, eng->last_items);
  }
  eng->release_retired();
  if (report) {
    report->kernel_ms = eng->kernel_ms - ms0;
    report->candidates = eng->candidates - c0;
    report->launches = eng->launches - l0;
    report->work_items = max_items;
  }
}

// Autotuner: times the job's applicable kernel variants on probes of its own candidates and,
// for the dense kernel of a one-problem job, its launch parameters: the alternative launch
// shapes of the tile bound (tile groups, ring depth), one pass fewer per wavefront round
// (candidates per round), and then the single-job launch size (candidates per launch: the
// default and 4x and 16x of it, as long as a launch stays near 50 ms).  Each configuration is
// run once to warm up and once timed over two of its launches (wall clock of the run, host
// merges included); the fastest is kept in the job, an alternative of the same variant only
// when it is more than 1% faster.  Results do not depend on the choice: candidate t is the same
// configuration in every kernel and shape, and records merge in candidate order.  The probe
// runs overwrite the job's results.
void autotune(sqsg_job *job, int *chosen, double *rate) {
  ensure_device(job);  // analyses the job's kernel applicability too
  struct Config {
    int variant;
    uint32_t alt, passes;
  };
  std::vector<Config> configs{{SQSG_KERNEL_SPARSE_WAVE, 0, 0}};
  if (job->lane_ok) configs.push_back({SQSG_KERNEL_SPARSE_LANE, 0, 0});
  const uint64_t saved_it = job->iterations, saved_launch = job->launch_size;
  const int saved_kernel = job->kernel;
  const uint32_t saved_alt = job->dense_alt, saved_passes = job->dense_passes;
  const uint64_t saved_tuned = job->tuned_launch;
  auto restore = [&] {
    job->iterations = saved_it, job->launch_size = saved_launch, job->kernel = saved_kernel;
    job->dense_alt = saved_alt, job->dense_passes = saved_passes;
    job->tuned_launch = saved_tuned;
  };
  job->dense_alt = 0, job->dense_passes = 0, job->tuned_launch = 0;
  if (job->dense_ok) {
    configs.push_back({SQSG_KERNEL_DENSE_MFMA, 0, 0});
    if (job->problems.size() == 1) {
      const uint32_t TT = dense_tile_bound_of(job);
      for (uint32_t a = 1; a <= dense_alt_shapes(TT); ++a)
        configs.push_back({SQSG_KERNEL_DENSE_MFMA, a, 0});
      uint32_t blocks = 0;
      const RunParams p = make_run_params(job, SQSG_KERNEL_DENSE_MFMA, 1, &blocks);
      const uint32_t npass = p.gen_group / p.dense_cw;
      if (npass > 1) configs.push_back({SQSG_KERNEL_DENSE_MFMA, 0, npass - 1});
    }
  }
  if (paired_ok(job)) configs.push_back({SQSG_KERNEL_DENSE_PAIRED, 0, 0});
  // each configuration runs with the launch size a run of it would use (the caller's, or
  // the variant's default) over about two of those launches
  auto launch_of = [&]() {
    return saved_launch ? saved_launch : default_launch_size(job);
  };
  auto timed = [&](uint64_t iterations, uint64_t launch) {
    job->iterations = iterations;
    job->launch_size = launch;
    const auto t0 = std::chrono::steady_clock::now();
    run_job(job, nullptr, nullptr);
    const double s =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return static_cast<double>(job->scored) / std::max(s, 1e-9);
  };
  Config best = configs[0];
  double best_rate = 0;
  try {
    for (const Config &c : configs) {
      job->kernel = c.variant, job->dense_alt = c.alt, job->dense_passes = c.passes;
      const uint64_t launch = std::min(launch_of(), saved_it);
      const uint64_t probe = std::min<uint64_t>(2 * launch, saved_it);
      (void)timed(probe, launch);  // warm-up (module load, first touch)
      const double r = timed(probe, launch);
      // an alternative of the same variant must win by more than the timing noise
      const bool same_variant = c.variant == best.variant && (c.alt || c.passes);
      if (r > best_rate * (same_variant ? 1.01 : 1.0)) best_rate = r, best = c;
    }
    job->kernel = best.variant, job->dense_alt = best.alt, job->dense_passes = best.passes;
    if (!saved_launch && saved_it > default_launch_size(job)) {
      // launch size of single-job runs: each launch ends with its slowest wavefront and a
      // host merge, so longer launches waste less, up to ~50 ms per launch
      const uint64_t base = default_launch_size(job);
      double best_l = 0;
      uint64_t pick = 0;
      for (uint64_t f : {1, 4, 16}) {
        const uint64_t L = base * f;
        if (f > 1 && (L > saved_it / 2 || static_cast<double>(L) / best_rate > 0.05)) break;
        const double r = timed(std::min<uint64_t>(3 * L, saved_it), L);
        if (r > best_l * (f > 1 ? 1.01 : 1.0)) best_l = r, pick = f > 1 ? L : 0;
      }
      job->tuned_launch = pick;
      best_rate = std::max(best_rate, best_l);
    }
  } catch (...) {
    restore();
    throw;
  }
  job->iterations = saved_it;
  job->launch_size = saved_launch;
  reset_results(job);
  if (chosen) *chosen = best.variant;
  if (rate) *rate = best_rate;
}

// ---- serving ---------------------------------------------------------------------------

// The kernel variant a job runs in the batched form; its device inputs (pair lists, the
// dense tile stream) are uploaded here.
void prepare_for_batch(sqsg_job *job) {
  ensure_device(job);
  const int variant = resolve_kernel(job, job->kernel == SQSG_KERNEL_SPARSE_WAVE
                                              ? SQSG_KERNEL_AUTO : job->kernel);
  if (variant == SQSG_KERNEL_DENSE_MFMA || variant == SQSG_KERNEL_DENSE_PAIRED)
    ensure_dense(job);
}

struct Ticket {
  sqsg_job *job = nullptr;
  double t_submit = -1, t_admit = -1, t_done = -1;
  bool done = false;
  size_t footprint = 0;
  std::atomic<int> state{0};  // 0: being prepared, 1: ready, 2: preparation failed
  std::exception_ptr error;
};

}  // namespace sqsg

struct sqsg_server {
  sqsg_server_options opt{};
  std::unique_ptr<sqsg::BatchEngine> eng;
  std::deque<uint64_t> queue;          // submitted, not admitted (submission order)
  std::deque<sqsg::Ticket> tickets;    // stable addresses: the preparation thread holds some
  std::map<sqsg_job *, uint64_t> active;
  size_t resident_bytes = 0;
  std::chrono::steady_clock::time_point t0;
  // preparation thread: uploads submitted jobs' inputs while the steps run
  int device = 0;
  std::thread prep;
  std::mutex m;
  std::condition_variable cv;
  std::deque<sqsg::Ticket *> prep_queue;
  bool stopping = false;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  ~sqsg_server() {
    {
      std::lock_guard<std::mutex> lock(m);
      stopping = true;
    }
    cv.notify_all();
    if (prep.joinable()) prep.join();
  }
};

namespace sqsg {

namespace {

void prepare_loop(sqsg_server *s) {
  (void)hipSetDevice(s->device);
  for (;;) {
    Ticket *t = nullptr;
    {
      std::unique_lock<std::mutex> lock(s->m);
      s->cv.wait(lock, [s] { return s->stopping || !s->prep_queue.empty(); });
      if (s->stopping) return;
      t = s->prep_queue.front();
      s->prep_queue.pop_front();
    }
    try {
      {
        std::lock_guard<std::mutex> capture_lock(capture_mutex());
        prepare_for_batch(t->job);
      }
      t->state.store(1, std::memory_order_release);
    } catch (...) {
      t->error = std::current_exception();
      t->state.store(2, std::memory_order_release);
    }
  }
}

// device bytes a job holds while resident: its inputs and two record sinks
size_t admission_bytes(const sqsg_job *job, uint64_t slice) {
  size_t rec = 24 + 8 * job->problems.size() + job->stride;
  return job_footprint(job) +
         2 * rec * first_sink_records(job, uint64_t(device_cu_count()) * 8,
                                      slice ? slice : default_step_size(job));
}

}  // namespace

sqsg_server *server_create(const sqsg_server_options &o) {
  auto s = std::make_unique<sqsg_server>();
  s->opt = o;
  if (!s->opt.max_active) s->opt.max_active = 64;
  SQSG_HIP(hipGetDevice(&s->device));
  if (!s->opt.hbm_budget_bytes) {  // most of the device memory that is free now
    size_t free_b = 0, total_b = 0;
    SQSG_HIP(hipMemGetInfo(&free_b, &total_b));
    s->opt.hbm_budget_bytes = free_b / 10 * 8;
  }
  s->eng = std::make_unique<BatchEngine>(true);
  s->eng->probe_new_ = true;
  s->t0 = std::chrono::steady_clock::now();
  s->prep = std::thread(prepare_loop, s.get());
  return s.release();
}

uint64_t server_submit(sqsg_server *s, sqsg_job *job) {
  s->tickets.emplace_back();
  Ticket &t = s->tickets.back();
  t.job = job;
  t.t_submit = s->now();
  const uint64_t id = s->tickets.size() - 1;
  s->queue.push_back(id);
  {
    std::lock_guard<std::mutex> lock(s->m);
    s->prep_queue.push_back(&t);
  }
  s->cv.notify_one();
  return id;
}

// Admission by footprint (in submission order, skipping jobs still being prepared or too
// large for the memory left), then one persistent step over the resident jobs.
void server_step(sqsg_server *s, int *active, int *queued) {
  for (auto it = s->queue.begin();
       it != s->queue.end() && s->active.size() < s->opt.max_active;) {
    Ticket &t = s->tickets[*it];
    const int st = t.state.load(std::memory_order_acquire);
    if (st == 0) {
      ++it;
      continue;
    }
    if (st == 2) {  // reported once: the ticket is done and the error is raised here
      t.done = true;
      t.t_done = s->now();
      s->queue.erase(it);
      std::rethrow_exception(t.error);
    }
    if (s->opt.slice_candidates) t.job->launch_size = s->opt.slice_candidates;
    const size_t need = admission_bytes(t.job, t.job->launch_size);
    if (!s->active.empty() && s->resident_bytes + need > s->opt.hbm_budget_bytes) {
      ++it;
      continue;
    }
    t.footprint = s->eng->add(t.job);
    s->resident_bytes += t.footprint;
    t.t_admit = s->now();
    s->active[t.job] = *it;
    it = s->queue.erase(it);
  }
  if (!s->eng->idle()) {
    for (sqsg_job *j : s->eng->step()) {
      const uint64_t id = s->active.at(j);
      Ticket &t = s->tickets[id];
      t.done = true;
      t.t_done = s->now();
      s->resident_bytes -= std::min(s->resident_bytes, t.footprint);
      s->active.erase(j);
    }
  }
  if (active) *active = static_cast<int>(s->active.size());
  if (queued) *queued = static_cast<int>(s->queue.size());
}

void server_poll(sqsg_server *s, uint64_t id, int *done, double *ts, double *ta, double *td) {
  if (id >= s->tickets.size()) throw std::out_of_range("unknown ticket");
  const Ticket &t = s->tickets[id];
  if (done) *done = t.done ? 1 : 0;
  if (ts) *ts = t.t_submit;
  if (ta) *ta = t.t_admit;
  if (td) *td = t.t_done;
}

void server_destroy(sqsg_server *s) { delete s; }

}  // namespace sqsg
