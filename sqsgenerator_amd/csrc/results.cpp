// Result collection on the host: sqs_result_collection (core/results.h:42-99) with the
// optimizer's search threshold (sqs.h:199-206,421-438), applied to the job's candidate stream
// in candidate order, which is the order one thread of the reference walks its blocks in.
// The device records every candidate at or below the threshold as it stood before the launch
// (a superset of what the reference inserts, since the threshold only decreases); the host
// replays those records in candidate order with the reference's rules: insert at or below
// the keep-th distinct objective inserted so far, one set of configurations per objective,
// at most max_results_per_objective + 1 per level, never evict.  Results are then mapped back
// to the original site order and atomic numbers (postprocess_results, sqs.h:94-122) with the
// reported rank of sqs_result_wrapper::rank (results.h:163-169).
//
// Levels can hold tens of millions of ties (the 32-site A16B16 exhaustive search has
// 17,271,296 optimal configurations), so records are stored as flat per-level arrays,
// deduplicated by hashing and ranked with 128-bit arithmetic on all host cores.
#include <algorithm>
#include <cfloat>
#include <cstring>
#include <limits>
#include <thread>
#include <unordered_map>

#include "job.hpp"
#include "kernels.hpp"
#include "permutation.hpp"

namespace sqsg {

namespace {

// 64-bit words mixed by multiply and shift (collisions are resolved by comparing bytes)
uint64_t hash_bytes(const uint8_t *p, size_t n) {
  uint64_t h = 0x9e3779b97f4a7c15ull ^ n;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, p + i, 8);
    h = (h ^ w) * 0xbf58476d1ce4e5b9ull;
    h ^= h >> 31;
  }
  uint64_t w = 0;
  std::memcpy(&w, p + i, n - i);
  h = (h ^ w) * 0x94d049bb133111ebull;
  return h ^ (h >> 29);
}

// 128-bit restatement of rank_permutation (permutation.h:34-50); exact whenever
// multinomial * n < 2^128, which the job checks once.
u128 rank_u128(const uint8_t *conf, size_t n) {
  uint32_t hist[256] = {0};
  bool seen[256] = {false};
  uint32_t num_species = 0;
  for (size_t i = 0; i < n; ++i)
    if (!seen[conf[i]]) seen[conf[i]] = true, ++num_species;
  u128 rank = 1, suffix = 1;
  for (size_t i = 0; i < n; ++i) {
    const uint8_t x = conf[(n - 1) - i];
    ++hist[x];
    for (uint32_t j = 0; j < num_species && j < x; ++j)
      if (hist[j]) rank += suffix * hist[j] / hist[x];
    suffix = suffix * (i + 1) / hist[x];
  }
  return rank;
}

template <class Fn> void parallel_for(uint64_t n, Fn &&fn) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096 || hw == 1) {
    fn(uint64_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const uint64_t chunk = (n + hw - 1) / hw;
  for (unsigned t = 0; t < hw; ++t) {
    const uint64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a < b) th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

void reset_results(sqsg_job *job) {
  job->coll.clear();
  job->coll.search = DBL_MAX;  // _search_objective starts at numeric_limits<T>::max()
  job->res.clear();
}

uint64_t block_of(const sqsg_job *job, uint64_t t) {
  if (t < job->blk_split) return t / (job->blk_size + 1);
  const uint64_t wide = job->blk_split / (job->blk_size + 1);
  return wide + (t - job->blk_split) / job->blk_size;
}

unsigned long long search_key(const sqsg_job *job) { return objective_key(job->coll.search); }

namespace {

// One range's records in candidate order (the worker loop of sqs.h:384-441 on one thread).
// In interact mode a candidate skipped for max_results_per_objective also skips the shuffle
// or the lexicographic successor that follows it (the `continue` at sqs.h:425), so the rest
// of its block re-scores that same configuration, which is skipped again: nothing else of
// that block is inserted.
void merge_in_order(sqsg_job *job, const HostRecords &r) {
  Collection &C = job->coll;
  const uint32_t nprob = static_cast<uint32_t>(job->problems.size());
  const uint32_t stride = job->stride;
  const bool systematic = job->iteration_mode == SQSG_ITERATION_SYSTEMATIC;
  const bool stalls = job->sublattice_mode == SQSG_SUBLATTICE_INTERACT;
  const RecordBlock &B = *r.block;
  for (const auto &[t, k] : r.order) {
    const double o = B.obj[k];
    if (!(o <= C.search)) continue;
    if (job->max_results >= 0 && C.stalled_block != ~uint64_t(0) &&
        block_of(job, t) == C.stalled_block)
      continue;
    auto it = C.levels.find(o);
    if (job->max_results >= 0 && it != C.levels.end() &&
        it->second.size() > static_cast<size_t>(job->max_results)) {
      if (stalls) C.stalled_block = block_of(job, t);
      continue;
    }
    const uint8_t *sp = B.species.data() + size_t(k) * stride;
    const bool fresh = it == C.levels.end();
    if (fresh) it = C.levels.emplace(o, LevelRecords{}).first;
    LevelRecords &lv = it->second;
    // a configuration repeats only in Monte Carlo runs
    if (!systematic && !lv.insert_unique(sp, stride, hash_bytes(sp, stride))) continue;
    lv.cand.push_back(t);
    lv.sub.insert(lv.sub.end(), B.sub.begin() + size_t(k) * nprob,
                  B.sub.begin() + size_t(k + 1) * nprob);
    lv.species.insert(lv.species.end(), sp, sp + stride);
    if (o < job->stats.best_objective) {
      job->stats.best_objective = o;
      job->stats.best_rank = t;
    }
    // update_objectives({objective, nth_best(keep)}) (sqs.h:435, 203-206)
    if (fresh && C.levels.size() > job->keep) {
      const double nb = std::next(C.levels.begin(), static_cast<long>(job->keep))->first;
      if (nb < C.search) C.search = nb;
    }
  }
}

}  // namespace

bool LevelRecords::insert_unique(const uint8_t *sp, uint32_t stride, uint64_t h) {
  if (2 * (size() + 1) > seen.size()) {  // grow (load at most one half), rehash
    size_t cap = 64;
    while (cap < 4 * (size() + 1)) cap *= 2;  // a power of two: slots are hash & (cap - 1)
    std::vector<std::pair<uint64_t, uint32_t>> old(cap, {0, 0});
    old.swap(seen);
    const size_t mask = seen.size() - 1;
    for (const auto &e : old)
      if (e.second) {
        size_t k = e.first & mask;
        while (seen[k].second) k = (k + 1) & mask;
        seen[k] = e;
      }
  }
  const size_t mask = seen.size() - 1;
  size_t k = h & mask;
  for (; seen[k].second; k = (k + 1) & mask)
    if (seen[k].first == h &&
        std::memcmp(species.data() + size_t(seen[k].second - 1) * stride, sp, stride) == 0)
      return false;
  seen[k] = {h, static_cast<uint32_t>(size() + 1)};
  return true;
}

unsigned long long merge_records(sqsg_job *job, HostRecords &&recs) {
  Collection &C = job->coll;
  if (recs.begin < C.merged) throw std::logic_error("candidate range merged twice");
  const uint64_t b = recs.begin;
  C.pending.emplace(b, std::move(recs));
  while (!C.pending.empty() && C.pending.begin()->first == C.merged) {
    HostRecords &r = C.pending.begin()->second;
    merge_in_order(job, r);
    C.merged = r.end;
    C.pending.erase(C.pending.begin());
  }
  return search_key(job);
}

void finalize_results(sqsg_job *job, bool eager_sro) {
  ResultSet &R = job->res;
  R.clear();
  if (!job->coll.pending.empty())
    throw std::logic_error("records of a candidate range were never merged");
  R.level_start.push_back(0);
  for (auto &[obj, lv] : job->coll.levels) {
    R.cand.insert(R.cand.end(), lv.cand.begin(), lv.cand.end());
    R.sub.insert(R.sub.end(), lv.sub.begin(), lv.sub.end());
    R.species.insert(R.species.end(), lv.species.begin(), lv.species.end());
    R.level_obj.push_back(obj);
    R.level_start.push_back(R.cand.size());
    lv = LevelRecords{};  // release as we go: packs can hold tens of millions of ties
  }
  job->coll.clear();
  job->coll.search = DBL_MAX;
  // SRO cubes are re-scored on the device, now or on first request
  if (eager_sro) cache_sro(job);
}

double result_objective(const sqsg_job *job, uint64_t r) {
  const auto &ls = job->res.level_start;
  const size_t l = std::upper_bound(ls.begin(), ls.end(), r) - ls.begin() - 1;
  return job->res.level_obj[l];
}

void result_species_z(const sqsg_job *job, uint64_t r, uint8_t *out) {
  const uint8_t *packed = job->res.species.data() + r * job->stride;
  std::memcpy(out, job->structure_z.data(), job->num_sites);
  if (job->sublattice_mode == SQSG_SUBLATTICE_INTERACT) {
    const Problem &P = job->problems[0];
    for (uint32_t k = 0; k < P.N; ++k) out[P.sort_order[k]] = P.zmap[packed[k]];
    return;
  }
  size_t o = 0;
  for (const auto &P : job->problems) {
    for (uint32_t k = 0; k < P.N; ++k) out[P.sites[k]] = P.zmap[packed[o + k]];
    o += P.N;
  }
}

namespace {

// packed original-order configurations whose ranks are reported, one per problem
template <class Fn> void for_each_ranked(const sqsg_job *job, uint64_t r, Fn &&fn) {
  const uint8_t *packed = job->res.species.data() + r * job->stride;
  if (job->sublattice_mode == SQSG_SUBLATTICE_INTERACT) {
    const Problem &P = job->problems[0];
    std::vector<uint8_t> orig(P.N);
    for (uint32_t k = 0; k < P.N; ++k) orig[P.sort_order[k]] = packed[k];
    fn(orig.data(), size_t(P.N));
    return;
  }
  size_t o = 0;
  for (const auto &P : job->problems) {
    fn(packed + o, size_t(P.N));
    o += P.N;
  }
}

}  // namespace

std::string result_rank(const sqsg_job *job, uint64_t r) {
  std::string out;
  for_each_ranked(job, r, [&](const uint8_t *c, size_t n) {
    if (!out.empty()) out += ",";
    out += job->rank_fast ? u128_str(rank_u128(c, n)) : rank_permutation(c, n).str();
  });
  return out;
}

bool result_rank_u64(const sqsg_job *job, uint64_t r, uint64_t *rank) {
  if (job->sublattice_mode != SQSG_SUBLATTICE_INTERACT || !job->rank_fast) return false;
  bool ok = false;
  for_each_ranked(job, r, [&](const uint8_t *c, size_t n) {
    const u128 v = rank_u128(c, n);
    ok = v <= std::numeric_limits<uint64_t>::max();
    *rank = ok ? static_cast<uint64_t>(v) : std::numeric_limits<uint64_t>::max();
  });
  return ok;
}

namespace {

// SRO cubes of results [first, first + count), re-scored on the device; each re-scored
// objective must equal the recorded one
void rescore_sro(sqsg_job *job, uint64_t first, uint64_t count, double *out) {
  const ResultSet &R = job->res;
  const uint64_t batch = uint64_t(1) << 20;
  for (uint64_t b0 = 0; b0 < count; b0 += batch) {
    const uint64_t B = std::min(batch, count - b0);
    size_t off = 0, cube_off = 0;
    for (uint32_t q = 0; q < job->problems.size(); ++q) {
      const Problem &P = job->problems[q];
      const size_t nb = size_t(P.S) * P.K * P.K;
      std::vector<uint8_t> sp(B * P.N);
      for (uint64_t k = 0; k < B; ++k)
        std::memcpy(&sp[k * P.N], R.species.data() + (first + b0 + k) * job->stride + off, P.N);
      DevBuf d_sp(sp.size()), d_obj(B * sizeof(double)), d_sro(B * nb * sizeof(double));
      SQSG_HIP(hipMemcpy(d_sp.as<void>(), sp.data(), sp.size(), hipMemcpyHostToDevice));
      score_batch(job, q, d_sp.as<uint8_t>(), static_cast<uint32_t>(B), nullptr,
                  d_obj.as<double>(), d_sro.as<double>(), SQSG_KERNEL_SPARSE_WAVE, job_stream(job));
      std::vector<double> obj(B), sro(B * nb);
      SQSG_HIP(hipMemcpyAsync(obj.data(), d_obj.as<void>(), B * sizeof(double),
                              hipMemcpyDeviceToHost, job_stream(job)));
      SQSG_HIP(hipMemcpyAsync(sro.data(), d_sro.as<void>(), sro.size() * sizeof(double),
                              hipMemcpyDeviceToHost, job_stream(job)));
      SQSG_HIP(hipStreamSynchronize(job_stream(job)));
      const uint32_t nprob = static_cast<uint32_t>(job->problems.size());
      for (uint64_t k = 0; k < B; ++k) {
        if (obj[k] != R.sub[(first + b0 + k) * nprob + q])
          throw DeviceError("re-scored objective differs from the recorded one");
        std::memcpy(out + (b0 + k) * job->cube_total + cube_off, &sro[k * nb],
                    nb * sizeof(double));
      }
      off += P.N;
      cube_off += nb;
    }
  }
}

}  // namespace

void cache_sro(sqsg_job *job) {
  ResultSet &R = job->res;
  if (!R.sro.empty() || !R.size() || R.size() > (uint64_t(1) << 20)) return;
  std::vector<double> sro(R.size() * job->cube_total);
  rescore_sro(job, 0, R.size(), sro.data());
  R.sro = std::move(sro);
}

void result_sro(sqsg_job *job, uint64_t first, uint64_t count, double *out) {
  const ResultSet &R = job->res;
  if (first + count > R.size()) throw std::out_of_range("result index out of range");
  cache_sro(job);  // packs of ordinary size: every cube at once, kept
  if (!R.sro.empty()) {
    std::memcpy(out, R.sro.data() + first * job->cube_total,
                count * job->cube_total * sizeof(double));
    return;
  }
  rescore_sro(job, first, count, out);
}

}  // namespace sqsg

namespace sqsg {

// objectives, candidates, original-order species and u64 ranks of results
// [first, first + count); split-mode ranks and ranks beyond 64 bits read UINT64_MAX
void result_bulk(sqsg_job *job, uint64_t first, uint64_t count, double *objective,
                 uint64_t *candidate, uint8_t *species_z, uint64_t *rank) {
  const ResultSet &R = job->res;
  if (first + count > R.size()) throw std::out_of_range("result index out of range");
  parallel_for(count, [&](uint64_t a, uint64_t b) {
    for (uint64_t k = a; k < b; ++k) {
      const uint64_t r = first + k;
      if (objective) objective[k] = result_objective(job, r);
      if (candidate) candidate[k] = R.cand[r];
      if (species_z) result_species_z(job, r, species_z + k * job->num_sites);
      if (rank && !result_rank_u64(job, r, rank + k)) rank[k] = UINT64_MAX;
    }
  });
}
}  // namespace sqsg
