// Batched persistent execution and continuous-batching service.
//
// BatchEngine keeps a set of resident jobs (their inputs, thresholds and record sinks stay
// in device memory) and advances all of them with one persistent launch per step: each
// step gives every active job its next slice of candidates as one work item, and the
// persistent wavefronts claim the items one segment at a time through per-item counters.
// Jobs are grouped by kernel variant and its specialisation (largest species count for the
// lane kernel, tile bound for the dense MFMA kernel); each group has fixed device tables
// and a captured graph {counter reset kernel, persistent kernel} that is replayed every
// step.  Between steps the host drains the record sinks, tightens each job's threshold and
// admits or retires jobs, which is what the serving loop builds on.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>

#include "job.hpp"
#include "kernels.hpp"

namespace sqsg {

namespace {

constexpr uint32_t kGroupJobs = 256;      // resident jobs per species-count group
constexpr uint32_t kGroupItems = kGroupJobs;  // one work item per resident job and step

struct Group;

struct Slot {
  sqsg_job *job = nullptr;
  Group *group = nullptr;
  RunParams p{};
  RecordSink sink;
  DevBuf thr, count, scored;
  uint64_t next = 0;         // first candidate of the next slice
  uint64_t slice = 0;        // candidates per step
  uint64_t chunk = 0;        // candidates a wavefront claims at once (one segment)
  uint64_t issued_end = 0;   // end of the slice issued in the current step
  unsigned int h_count = 0;  // host copies of the device counters after a step
  unsigned long long h_scored = 0, h_thr = 0;
  size_t footprint = 0;      // device bytes held
  bool done = false;
};

struct Group {
  int variant = 0;
  uint32_t key = 0;  // lane: largest species count; dense: tile bound
  DevBuf d_jobs, d_items, d_counts, d_claims;
  hipGraphExec_t exec = nullptr;
  uint32_t blocks = 0;     // of the current step's launch
  BatchParams bp{};
  std::vector<RunParams> h_table;  // host sources of the step's uploads
  std::vector<WorkItem> h_items;
  uint32_t h_counts[2] = {0, 0};
  uint32_t graph_wave_lds = 0, graph_blocks = 0;
  std::vector<Slot *> members;
  ~Group() {
    if (exec) (void)hipGraphExecDestroy(exec);
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
    SQSG_HIP(hipEventCreate(&ev0_));
    SQSG_HIP(hipEventCreate(&ev1_));
  }
  ~BatchEngine() {
    groups_.clear();
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
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
    if (variant == SQSG_KERNEL_DENSE_MFMA) ensure_dense(job);
    auto s = std::make_unique<Slot>();
    s->job = job;
    uint32_t blocks = 0;
    s->p = make_run_params(job, variant, 1, &blocks);
    s->slice = job->launch_size ? job->launch_size : default_launch_size(job);
    uint64_t work = 0;
    for (const auto &P : job->problems) work += P.pi.size() + 64;
    const uint64_t gen = s->p.gen_group;
    if (s->p.systematic) {
      s->p.rounds = static_cast<uint32_t>(std::clamp<uint64_t>(
          (uint64_t(1) << 21) / (gen * std::max<uint64_t>(work, 1)), 1, 256));
    } else {
      s->p.rounds = 1;
    }
    s->chunk = gen * s->p.rounds;
    s->slice = std::max(s->slice, s->chunk);
    s->p.rec_stride = job->stride;
    s->sink.ensure(static_cast<uint32_t>(std::min<uint64_t>(s->slice, 1u << 14)), job->stride,
                   s->p.nprob);
    s->thr.alloc(sizeof(unsigned long long));
    s->count.alloc(sizeof(unsigned int));
    s->scored.alloc(sizeof(unsigned long long));
    const unsigned long long inf = objective_key(std::numeric_limits<double>::infinity());
    SQSG_HIP(hipMemcpy(s->thr.as<void>(), &inf, sizeof(inf), hipMemcpyHostToDevice));
    s->p.thr_key = s->thr.as<unsigned long long>();
    s->p.rec_count = s->count.as<unsigned int>();
    s->p.scored = s->scored.as<unsigned long long>();
    s->footprint = job_footprint(job) + s->sink.cap * (24 + job->stride);
    Group &g = group(variant, variant == SQSG_KERNEL_DENSE_MFMA ? dense_tile_bound_of(job)
                                                                : job->kmax);
    if (g.members.size() >= kGroupJobs) throw std::out_of_range("too many resident jobs");
    g.members.push_back(s.get());
    s->group = &g;
    const size_t fp = s->footprint;
    slots_.push_back(std::move(s));
    return fp;
  }

  bool idle() const { return slots_.empty(); }
  size_t resident() const { return slots_.size(); }

  // One persistent launch over every active job; finished jobs are finalised and returned.
  std::vector<sqsg_job *> step() {
    std::vector<sqsg_job *> finished;
    if (slots_.empty()) return finished;
    const auto t0 = std::chrono::steady_clock::now();
    // tables first, then the launches bracketed by the step's events
    std::vector<Group *> live;
    for (auto &[k, g] : groups_)
      if (!g->members.empty()) live.push_back(prepare_group(*g));
    SQSG_HIP(hipEventRecord(ev0_, stream_));
    for (Group *g : live) launch_group(*g);
    SQSG_HIP(hipEventRecord(ev1_, stream_));
    SQSG_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    SQSG_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_ms += ms;
    ++launches;
    // every slot's counters and threshold in one round trip
    for (auto &s : slots_) {
      if (s->issued_end == s->next) continue;
      SQSG_HIP(hipMemcpyAsync(&s->h_count, s->count.as<void>(), sizeof(s->h_count),
                              hipMemcpyDeviceToHost, stream_));
      SQSG_HIP(hipMemcpyAsync(&s->h_scored, s->scored.as<void>(), sizeof(s->h_scored),
                              hipMemcpyDeviceToHost, stream_));
      SQSG_HIP(hipMemcpyAsync(&s->h_thr, s->thr.as<void>(), sizeof(s->h_thr),
                              hipMemcpyDeviceToHost, stream_));
    }
    SQSG_HIP(hipStreamSynchronize(stream_));
    for (auto &s : slots_) {
      sqsg_job *job = s->job;
      if (s->issued_end == s->next) {
        if (s->next >= job->iterations || job->cancel.load()) s->done = true;
        continue;
      }
      const unsigned int count = s->h_count;
      const unsigned long long scored = s->h_scored;
      job->launches++;
      job->kernel_ms += ms;  // shared launch: every resident job saw the whole step
      if (count > s->sink.cap) {  // widen and issue the same slice again next step
        s->sink.ensure(count + count / 4, job->stride, s->p.nprob);
        s->footprint = job_footprint(job) + s->sink.cap * (24 + job->stride);
        continue;
      }
      if (scored != s->issued_end - s->next)
        throw DeviceError("persistent launch scored an unexpected candidate count");
      job->scored += scored;
      candidates += scored;
      s->sink.drain(job, count, s->p.nprob, stream_);
      // keep the device threshold at the tighter of its own bound and the exact host one
      // (ordered before the next launch on the same stream; h_thr outlives the copy)
      const unsigned long long hk = objective_key(job->vkeep);
      if (hk < s->h_thr) {
        s->h_thr = hk;
        SQSG_HIP(hipMemcpyAsync(s->thr.as<void>(), &s->h_thr, sizeof(s->h_thr),
                                hipMemcpyHostToDevice, stream_));
      }
      s->next = s->issued_end;
      job->stats.finished = job->scored;
      if (s->next >= job->iterations || job->cancel.load()) s->done = true;
    }
    const auto t1 = std::chrono::steady_clock::now();
    for (auto it = slots_.begin(); it != slots_.end();) {
      if (!(*it)->done) {
        ++it;
        continue;
      }
      sqsg_job *job = (*it)->job;
      finalize_results(job);
      job->stats.timing_loop_ns +=
          std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
      job->stats.timing_total_ns = job->stats.timing_loop_ns;
      Group &g = *(*it)->group;
      g.members.erase(std::find(g.members.begin(), g.members.end(), it->get()));
      finished.push_back(job);
      it = slots_.erase(it);
    }
    for (auto &s : slots_)
      s->job->stats.timing_loop_ns +=
          std::chrono::duration_cast<std::chrono::nanoseconds>(t1 - t0).count();
    return finished;
  }

  double kernel_ms = 0;
  uint64_t candidates = 0, launches = 0;
  uint32_t last_items = 0;

 private:
  Group &group(int variant, uint32_t key) {
    auto &g = groups_[(static_cast<uint32_t>(variant) << 8) | key];
    if (!g) {
      g = std::make_unique<Group>();
      g->variant = variant;
      g->key = key;
      g->d_jobs.alloc(kGroupJobs * sizeof(RunParams));
      g->d_items.alloc(kGroupItems * sizeof(WorkItem));
      g->d_counts.alloc(2 * sizeof(uint32_t));
      g->d_claims.alloc(kGroupItems * sizeof(unsigned int));
    }
    return *g;
  }

  // uploads the group's job table and work items for this step
  Group *prepare_group(Group &g) {
    std::vector<RunParams> &table = g.h_table;
    std::vector<WorkItem> &items = g.h_items;
    table.resize(g.members.size());
    items.clear();
    uint32_t wave_lds = 0;
    for (uint32_t k = 0; k < g.members.size(); ++k) {
      Slot *s = g.members[k];
      s->sink.bind(s->p);
      table[k] = s->p;
      wave_lds = std::max(wave_lds, s->p.wave_lds);
      const uint64_t end = std::min(s->job->iterations, s->next + s->slice);
      s->issued_end = s->job->cancel.load() ? s->next : end;
      if (s->issued_end > s->next)
        items.push_back(WorkItem{k, static_cast<uint32_t>(s->chunk), s->next, s->issued_end});
    }
    last_items = static_cast<uint32_t>(items.size());
    const uint32_t n = last_items;
    SQSG_HIP(hipMemcpyAsync(g.d_jobs.as<void>(), table.data(), table.size() * sizeof(RunParams),
                            hipMemcpyHostToDevice, stream_));
    if (n)
      SQSG_HIP(hipMemcpyAsync(g.d_items.as<void>(), items.data(), n * sizeof(WorkItem),
                              hipMemcpyHostToDevice, stream_));
    g.h_counts[0] = n;
    g.h_counts[1] = static_cast<uint32_t>(g.members.size());
    // the host tables stay untouched until the step's end synchronisation
    SQSG_HIP(hipMemcpyAsync(g.d_counts.as<void>(), g.h_counts, sizeof(g.h_counts),
                            hipMemcpyHostToDevice, stream_));
    const size_t block_lds = wave_lds;
    const bool dense = g.variant == SQSG_KERNEL_DENSE_MFMA;
    uint32_t per_cu = static_cast<uint32_t>((160 * 1024) / std::max<size_t>(block_lds, 1));
    per_cu = std::max(1u, std::min(per_cu, 16u));
    if (dense) per_cu = dense_waves_per_cu(g.key, block_lds, true);
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
    g.blocks = blocks;
    SQSG_HIP(dense ? prepare_batch_dense(g.key, block_lds) : prepare_batch_lane(g.key, block_lds));
    if (use_graph_ && (!g.exec || g.graph_wave_lds != wave_lds || g.graph_blocks != blocks)) {
      // the graph covers every table row (num_jobs is a kernel argument of the reset)
      if (g.exec) SQSG_HIP(hipGraphExecDestroy(g.exec));
      g.exec = nullptr;
      hipGraph_t graph = nullptr;
      SQSG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      const hipError_t e1 = launch_batch_reset(bp, kGroupJobs, stream_);
      const hipError_t e2 = launch_kernel(g);
      SQSG_HIP(hipStreamEndCapture(stream_, &graph));
      SQSG_HIP(e1);
      SQSG_HIP(e2);
      const hipError_t e3 = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      SQSG_HIP(e3);
      g.graph_wave_lds = wave_lds;
      g.graph_blocks = blocks;
    }
    return &g;
  }

  hipError_t launch_kernel(const Group &g) {
    return g.variant == SQSG_KERNEL_DENSE_MFMA
               ? launch_batch_dense(g.bp, g.key, g.blocks, stream_)
               : launch_batch_lane(g.bp, g.key, g.blocks, 1, stream_);
  }

  // the step's launches only (tables, attributes and graphs are prepared beforehand)
  void launch_group(Group &g) {
    if (!use_graph_) {
      SQSG_HIP(launch_batch_reset(g.bp, kGroupJobs, stream_));
      SQSG_HIP(launch_kernel(g));
      return;
    }
    SQSG_HIP(hipGraphLaunch(g.exec, stream_));
  }

  bool use_graph_;
  hipStream_t stream_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::map<uint32_t, std::unique_ptr<Group>> groups_;
};

void batch_run(sqsg_job **jobs, uint32_t n, int use_graph, sqsg_batch_report *report) {
  BatchEngine eng(use_graph != 0);
  for (uint32_t k = 0; k < n; ++k) eng.add(jobs[k]);
  uint32_t max_items = 0;
  while (!eng.idle()) {
    eng.step();
    max_items = std::max(max_items, eng.last_items);
  }
  if (report) {
    report->kernel_ms = eng.kernel_ms;
    report->candidates = eng.candidates;
    report->launches = eng.launches;
    report->work_items = max_items;
  }
}

// Times every applicable kernel variant on a short launch of the job and keeps the fastest.
// The probe runs overwrite the job's results.
void autotune(sqsg_job *job, int *chosen, double *rate) {
  ensure_device(job);
  std::vector<int> variants{SQSG_KERNEL_SPARSE_WAVE};
  if (job->lane_ok) variants.push_back(SQSG_KERNEL_SPARSE_LANE);
  if (job->dense_ok) variants.push_back(SQSG_KERNEL_DENSE_MFMA);
  const uint64_t saved_it = job->iterations, saved_launch = job->launch_size;
  const int saved_kernel = job->kernel;
  uint64_t work = 0;
  for (const auto &P : job->problems) work += P.pi.size() + 64;
  const uint64_t probe = std::clamp<uint64_t>((uint64_t(1) << 31) / work, 4096, saved_it);
  int best = SQSG_KERNEL_SPARSE_WAVE;
  double best_rate = 0;
  try {
    job->iterations = probe;
    job->launch_size = probe;
    for (int v : variants) {
      job->kernel = v;
      run_job(job, nullptr, nullptr);  // warm-up (module load, first-touch)
      run_job(job, nullptr, nullptr);
      const double r = job->scored / (job->kernel_ms * 1e-3);
      if (r > best_rate) best_rate = r, best = v;
    }
  } catch (...) {
    job->iterations = saved_it, job->launch_size = saved_launch, job->kernel = saved_kernel;
    throw;
  }
  job->iterations = saved_it;
  job->launch_size = saved_launch;
  job->kernel = best;
  reset_results(job);
  if (chosen) *chosen = best;
  if (rate) *rate = best_rate;
}

// ---- serving ---------------------------------------------------------------------------

struct Ticket {
  sqsg_job *job;
  double t_submit = -1, t_admit = -1, t_done = -1;
  bool done = false;
  size_t footprint = 0;
};

}  // namespace sqsg

struct sqsg_server {
  sqsg_server_options opt{};
  std::unique_ptr<sqsg::BatchEngine> eng;
  std::deque<uint64_t> queue;
  std::vector<sqsg::Ticket> tickets;
  std::map<sqsg_job *, uint64_t> active;
  size_t resident_bytes = 0;
  std::chrono::steady_clock::time_point t0;
  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace sqsg {

sqsg_server *server_create(const sqsg_server_options &o) {
  auto s = std::make_unique<sqsg_server>();
  s->opt = o;
  if (!s->opt.max_active) s->opt.max_active = 64;
  if (!s->opt.hbm_budget_bytes) s->opt.hbm_budget_bytes = size_t(8) << 30;
  s->eng = std::make_unique<BatchEngine>(true);
  s->t0 = std::chrono::steady_clock::now();
  return s.release();
}

uint64_t server_submit(sqsg_server *s, sqsg_job *job) {
  Ticket t;
  t.job = job;
  t.t_submit = s->now();
  s->tickets.push_back(t);
  const uint64_t id = s->tickets.size() - 1;
  s->queue.push_back(id);
  return id;
}

// Admission by footprint, then one persistent step over the resident jobs.
void server_step(sqsg_server *s, int *active, int *queued) {
  while (!s->queue.empty() && s->active.size() < s->opt.max_active) {
    Ticket &t = s->tickets[s->queue.front()];
    ensure_device(t.job);
    const size_t need = job_footprint(t.job) + (size_t(1) << 14) * (24 + t.job->stride);
    if (!s->active.empty() && s->resident_bytes + need > s->opt.hbm_budget_bytes) break;
    if (s->opt.slice_candidates) t.job->launch_size = s->opt.slice_candidates;
    t.footprint = s->eng->add(t.job);
    s->resident_bytes += t.footprint;
    t.t_admit = s->now();
    s->active[t.job] = s->queue.front();
    s->queue.pop_front();
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
