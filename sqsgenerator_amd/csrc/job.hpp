// Host-side job: optimization_config (include/sqsgen/core/optimization_config.h) plus
// the device-resident copies of its read-only inputs, the run state and the results.
#pragma once

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <map>
#include <memory>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/sqsgen_amd.h"
#include "device_params.hpp"
#include "geometry.hpp"

namespace sqsg {

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct CancelledError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what);
#define SQSG_HIP(call) ::sqsg::hip_check((call), #call)

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) {
      reset();
      p_ = o.p_, n_ = o.n_;
      o.p_ = nullptr, o.n_ = 0;
    }
    return *this;
  }
  void alloc(size_t bytes);
  void reset();
  template <class T> T *as() const { return static_cast<T *>(p_); }
  size_t size() const { return n_; }

 private:
  void *p_ = nullptr;
  size_t n_ = 0;
};

// One sublattice problem (interact: the whole structure; split: one sublattice).
struct Problem {
  uint32_t N = 0, K = 0, S = 0;
  bool single = false;                  // configuration<float>
  std::vector<uint32_t> pi, pj, ps;     // pairs in the reference's build order (i, then j)
  std::vector<uint8_t> species_packed;  // sorted-site order, packed indices
  std::vector<uint32_t> sort_order;     // sorted position -> site of the problem's structure
  std::vector<uint32_t> bounds;         // (lo, hi) shuffle windows
  std::vector<uint8_t> zmap;            // packed index -> Z
  std::vector<uint32_t> shell_ids;      // packed shell -> shell id
  std::vector<double> pref, pw, tgt;    // [S][K][K]; pw scaled by the shell weights
  std::vector<uint32_t> terms;          // objective terms in reference order
  std::vector<uint32_t> sites;          // problem site -> original structure site
  uint64_t seed = 0;
  uint32_t draws = 0;
  // systematic window
  uint32_t sys_nspecies = 0;
  uint32_t sys_hist[kMaxLocalSpecies] = {};
  uint8_t sys_rev[kMaxLocalSpecies] = {};
  // device copies
  DevBuf d_pairs, d_species0, d_pref, d_pw, d_tgt, d_terms, d_tparams;
  // lane-per-candidate layout (built for the job's field width F)
  std::vector<uint32_t> lane_stream, lane_off;
  uint32_t F = 0;
  DevBuf d_lane_stream, d_lane_off, d_species0F;
  void build_lane(uint32_t field_bits);
  // dense contraction: 32-site tiles of the upper shell adjacency (built on first use)
  uint32_t dT = 0;          // 32-site tiles
  uint32_t max_upper = 0;   // largest count of upper neighbours of one site in one shell
  uint32_t dtiles_n = 0;
  uint32_t dtiles_nonzero = 0;  // tiles holding at least one pair
  DevBuf d_dtiles, d_species0D;
  // one-hot columns of the dense search kernels (K, or K - 1: DevProblem::Kc) and the
  // per-shell, per-species neighbour-count sums that let the last species go without one
  uint32_t Kc = 0;
  std::vector<uint32_t> dsum;
  DevBuf d_dsum;
  void reduce_columns();
  void build_dense(uint32_t tile_bound);  // the kernel's tile bound (the job's largest)
  DevProblem dev() const;
  void upload_device();
  uint32_t row_stride() const { return (((N + 3) / 4) | 1u) * 4u; }
};

// One objective level of the collection: the set of configurations inserted at that
// objective (results.h:46-53), in candidate order.
struct LevelRecords {
  std::vector<uint64_t> cand;
  std::vector<double> sub;        // one objective per problem and record
  std::vector<uint8_t> species;   // sorted-site packed species, `stride` bytes per record
  // Monte Carlo: open-addressing set of the level's configurations (hash, record + 1)
  std::vector<std::pair<uint64_t, uint32_t>> seen;
  size_t size() const { return cand.size(); }
  // inserts the configuration at `sp` (stride bytes, hash h) unless the level holds it
  bool insert_unique(const uint8_t *sp, uint32_t stride, uint64_t h);
};

// One copy of a record sink's contents on the host.
struct RecordBlock {
  std::vector<uint64_t> cand;
  std::vector<double> obj, sub;
  std::vector<uint8_t> species;
};

// The device records of one candidate range [begin, end): (candidate, record) pairs of a
// block, in candidate order.
struct HostRecords {
  uint64_t begin = 0, end = 0;
  std::shared_ptr<const RecordBlock> block;
  std::vector<std::pair<uint64_t, uint32_t>> order;
};

// The reference's result collection and search threshold, applied to the job's candidate
// stream in candidate order (one thread's run of the worker loop, sqs.h:384-441): candidate
// t is inserted when its objective is at or below `search`, the keep-th distinct objective
// inserted before it (results.h:78-82; DBL_MAX while fewer exist, sqs.h:259-260), unless its
// level already holds more than max_results_per_objective configurations (sqs.h:423-426);
// levels are never evicted.  The device records a superset of the inserted candidates;
// ranges that come back out of order (a record sink that overflowed runs again) wait in
// `pending` until every earlier candidate has been merged.
struct Collection {
  std::map<double, LevelRecords> levels;
  double search = 0;
  uint64_t merged = 0;                       // candidates [0, merged) are merged
  uint64_t stalled_block = ~uint64_t(0);     // see merge_in_order
  std::map<uint64_t, HostRecords> pending;   // by range begin
  void clear() { *this = Collection(); }
};

// Final, deduplicated results in ascending objective order (the sqs_result_pack levels).
struct ResultSet {
  std::vector<double> level_obj;
  std::vector<uint64_t> level_start;  // levels + 1 offsets into the per-result arrays
  std::vector<uint64_t> cand;
  std::vector<double> sub;
  std::vector<uint8_t> species;
  std::vector<double> sro;            // cached SRO cubes (all problems) or empty
  size_t size() const { return cand.size(); }
  void clear() { *this = ResultSet(); }
};

}  // namespace sqsg

struct sqsg_job {
  int iteration_mode = 0, sublattice_mode = 0;
  bool single = false;  // configuration<float>: float geometry, weights and objectives
  std::vector<sqsg::Problem> problems;
  uint32_t num_sites = 0;
  std::vector<uint8_t> structure_z;  // applied composition, original order
  uint64_t iterations = 0, chunk_size = 0, keep = 1;
  int64_t max_results = -1;
  int kernel = SQSG_KERNEL_AUTO;
  bool prune = false;     // early exit of hopeless passes (dense kernel, pw >= 0)
  sqsg::DevBuf score_scratch;  // term values of explicit scoring split by shell
  uint64_t launch_size = 0;
  // launch parameters chosen by the autotuner for the dense search kernels of a job of one
  // problem: an alternative launch shape of the tile bound (0: the default), candidates per
  // wavefront round in passes of the shape's width (0: as many as its LDS budget holds), and
  // single-job candidates per launch (0: default_launch_size; an explicit launch size wins)
  uint32_t dense_alt = 0, dense_passes = 0;
  uint64_t tuned_launch = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool on_device = false;
  bool analyzed = false;  // kmax, lane_ok, dense_ok set
  uint32_t kmax = 0;      // largest species count over the problems
  bool lane_ok = false;   // the lane-per-candidate kernel applies
  bool dense_ok = false;  // the dense MFMA contraction applies
  std::atomic<bool> cancel{false};
  // run state
  // blocks of the reference's thread-pool split of [0, iterations) (sqs.h:461-466)
  uint64_t blk_size = 0, blk_split = 0;
  sqsg::Collection coll;
  uint32_t stride = 0;        // packed species bytes per record (sum of problem sites)
  size_t cube_total = 0;      // doubles of one result's SRO cubes
  bool rank_fast = false;     // reported ranks fit 128-bit arithmetic
  sqsg_statistics stats{};
  double kernel_ms = 0;
  uint64_t launches = 0, scored = 0;
  sqsg::ResultSet res;
  ~sqsg_job();
};

namespace sqsg {

// merges the records of a host copy of a sink, split by candidate range, into the job's
// collection in candidate order
void merge_block(sqsg_job *job, std::shared_ptr<const RecordBlock> blk,
                 const std::vector<std::pair<uint64_t, uint64_t>> &ranges);
// host copy of the first `count` records of a persistent step in a job's host stage
// (RunParams::stage layout, room for `cap` records)
std::shared_ptr<RecordBlock> stage_block(const sqsg_job *job, const uint8_t *stage,
                                         uint32_t cap, uint32_t count, uint32_t nprob);
// Device memory a run uses instead of allocating its own: control words and the record sink.
struct FixedScratch {
  void *base = nullptr;
  size_t bytes = 0;
};
sqsg_job *create_job(const sqsg_config &cfg);
// a run of the job; with `fixed`, its threshold, counters and record sink live in that
// memory and the run allocates no device memory (the job's inputs must be resident)
void run_job(sqsg_job *job, sqsg_progress_fn cb, void *user,
             const FixedScratch *fixed = nullptr);
// runs fn, translating exceptions into an SQSG_* status and the thread's last error text
int guarded(const std::function<void()> &fn);
void score_batch(sqsg_job *job, uint32_t problem, const uint8_t *d_species, uint32_t B,
                 int32_t *d_bonds, double *d_obj, double *d_sro, int kernel, hipStream_t s);
// helpers shared with the batched and serving drivers
// (search: with the job's tuned launch parameters; explicit scoring uses the defaults)
RunParams make_run_params(sqsg_job *job, int variant, uint32_t waves_per_block,
                          uint32_t *blocks, bool search = true);
// the tuned alternative dense shape a search of the job runs (0: the default shape)
uint32_t search_alt(const sqsg_job *job);
// Device record sink of one job: candidates at or below the threshold with their
// objectives and packed species.  Its arrays are its own allocations, or memory the caller
// lends it (place).
struct RecordSink {
  DevBuf cand_buf, obj_buf, sub_buf, species_buf;
  uint64_t *cand = nullptr;
  double *obj = nullptr, *sub = nullptr;
  uint8_t *species = nullptr;
  uint32_t cap = 0;
  void ensure(uint32_t n, uint32_t stride, uint32_t nprob);
  // carves a sink out of `bytes` of device memory at `base` (256-byte aligned); returns
  // its capacity in records
  uint32_t place(void *base, size_t bytes, uint32_t stride, uint32_t nprob);
  void bind(RunParams &p) const;
  // copies `count` records back to the host
  std::shared_ptr<RecordBlock> fetch(const sqsg_job *job, uint32_t count, uint32_t nprob,
                                     hipStream_t s) const;
  // copies `count` records of the candidate ranges back and merges them into the job's
  // collection in candidate order
  void drain(sqsg_job *job, uint32_t count, uint32_t nprob,
             const std::vector<std::pair<uint64_t, uint64_t>> &ranges, hipStream_t s);
};
// candidates per launch of a single-job run (sqsg_run) without an explicit launch size
uint64_t default_launch_size(const sqsg_job *job);
// candidates per persistent step of a batched or served job without an explicit size:
// about 2^34 pair visits, so that a job's threshold tightens after its first short steps
uint64_t default_step_size(const sqsg_job *job);
// the wavefront filter bounds the device records of the job (RunParams::wave_filter)
bool records_bounded(const sqsg_job *job);
// initial record sink size of a launch of `launch` candidates over `waves` wavefronts
uint32_t first_sink_records(const sqsg_job *job, uint64_t waves, uint64_t launch);
void begin_run(sqsg_job *job);
// kernel variant a run uses (SQSG_KERNEL_SPARSE_WAVE / _LANE / _DENSE_MFMA)
int resolve_kernel(const sqsg_job *job, int requested);
hipError_t launch_variant(const RunParams &p, const sqsg_job *job, int variant, uint32_t blocks,
                          uint32_t waves_per_block, hipStream_t s);
void reset_results(sqsg_job *job);
// hands the records of one candidate range to the collection (merged once every earlier
// range is); returns the device threshold key for the next launch
unsigned long long merge_records(sqsg_job *job, HostRecords &&recs);
unsigned long long search_key(const sqsg_job *job);
// block of the reference's thread-pool split that candidate t belongs to
uint64_t block_of(const sqsg_job *job, uint64_t t);
// flattens the collection into the result set; eager_sro re-scores the SRO cubes on the
// device now (else on first request)
void finalize_results(sqsg_job *job, bool eager_sro = true);
// accessors of finalized result r (original site order, atomic numbers)
void result_species_z(const sqsg_job *job, uint64_t r, uint8_t *out);
std::string result_rank(const sqsg_job *job, uint64_t r);
bool result_rank_u64(const sqsg_job *job, uint64_t r, uint64_t *rank);
void result_sro(sqsg_job *job, uint64_t first, uint64_t count, double *out);
// re-scores and keeps the SRO cubes of every result (packs of up to 2^20 results)
void cache_sro(sqsg_job *job);
double result_objective(const sqsg_job *job, uint64_t r);
void result_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *objective,
                 uint64_t *candidate, uint8_t *species_z, uint64_t *rank);
int device_cu_count();
void batch_run(sqsg_job **jobs, uint32_t n, int use_graph, sqsg_batch_report *report);
void autotune(sqsg_job *job, int *chosen, double *rate);
sqsg_server *server_create(const sqsg_server_options &o);
uint64_t server_submit(sqsg_server *s, sqsg_job *job);
void server_step(sqsg_server *s, int *active, int *queued);
void server_poll(sqsg_server *s, uint64_t id, int *done, double *ts, double *ta, double *td);
void server_destroy(sqsg_server *s);
void ensure_device(sqsg_job *job);
// the job's own stream (created on first use)
hipStream_t job_stream(sqsg_job *job);
void ensure_dense(sqsg_job *job);
// dense variant launch shape of a job: tile bound (kernel template) and 1-wave blocks
uint32_t dense_tile_bound_of(const sqsg_job *job);
// which kernel variants apply to the job (host data; kmax, lane_ok, dense_ok)
void analyze_job(sqsg_job *job);
// the paired dense form applies (Monte Carlo, a tile bound with a paired shape)
bool paired_ok(const sqsg_job *job);
// launch shape of the kernel a run of the job uses ("dense<TT,G,W,D>", "lane<K>", ...)
std::string kernel_shape(sqsg_job *job);

}  // namespace sqsg
