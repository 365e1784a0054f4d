// Result collection on the host: the deterministic part of sqs_result_collection
// (core/results.h:42-99).  The device appends every candidate at or below its running
// threshold; the host keeps the keep + 1 best distinct objective levels with all their
// ties, deduplicates configurations per level (one set per objective, results.h:46-53),
// applies max_results_per_objective (the reference inserts while the level holds at most
// `max` entries, sqs.h:423-426), and maps results back to the original site order and
// atomic numbers (postprocess_results, sqs.h:94-122) with the reported rank of
// sqs_result_wrapper::rank (results.h:163-169).
//
// Levels can hold tens of millions of ties (the 32-site A16B16 exhaustive search has
// 17,271,296 optimal configurations), so records are stored as flat per-level arrays,
// deduplicated by hashing and ranked with 128-bit arithmetic on all host cores.
#include <algorithm>
#include <cstring>
#include <limits>
#include <thread>
#include <unordered_map>

#include "job.hpp"
#include "permutation.hpp"

namespace sqsg {

namespace {

uint64_t hash_bytes(const uint8_t *p, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ p[i]) * 0x100000001b3ull;
  return h ^ (h >> 31);
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
  job->levels.clear();
  job->res.clear();
  job->vkeep = std::numeric_limits<double>::infinity();
}

void merge_records(sqsg_job *job, const uint64_t *cand, const double *obj, const double *sub,
                   const uint8_t *species, uint32_t count) {
  const uint32_t nprob = static_cast<uint32_t>(job->problems.size());
  const uint32_t stride = job->stride;
  const size_t L = job->keep + 1;
  for (uint32_t k = 0; k < count; ++k) {
    if (obj[k] > job->vkeep) continue;
    auto &lv = job->levels[obj[k]];
    lv.cand.push_back(cand[k]);
    lv.sub.insert(lv.sub.end(), sub + size_t(k) * nprob, sub + size_t(k + 1) * nprob);
    lv.species.insert(lv.species.end(), species + size_t(k) * stride,
                      species + size_t(k + 1) * stride);
    if (job->levels.size() > L) {
      job->levels.erase(std::prev(job->levels.end()));
      job->vkeep = std::prev(job->levels.end())->first;
    } else if (job->levels.size() == L) {
      job->vkeep = std::prev(job->levels.end())->first;
    }
  }
}

void finalize_results(sqsg_job *job) {
  ResultSet &R = job->res;
  R.clear();
  const uint32_t nprob = static_cast<uint32_t>(job->problems.size());
  const uint32_t stride = job->stride;
  const bool systematic = job->iteration_mode == SQSG_ITERATION_SYSTEMATIC;
  R.level_start.push_back(0);
  for (auto &[obj, lv] : job->levels) {
    const size_t n = lv.cand.size();
    std::vector<uint32_t> order(n);
    for (size_t k = 0; k < n; ++k) order[k] = static_cast<uint32_t>(k);
    std::sort(order.begin(), order.end(),
              [&](uint32_t a, uint32_t b) { return lv.cand[a] < lv.cand[b]; });
    // a candidate index is written by exactly one record; configurations repeat only in
    // MC (the same permutation drawn twice), never in systematic enumeration
    std::unordered_map<uint64_t, std::vector<uint32_t>> seen;
    if (!systematic) seen.reserve(n * 2);
    size_t kept = 0;
    for (uint32_t k : order) {
      if (job->max_results >= 0 && kept > static_cast<size_t>(job->max_results)) break;
      const uint8_t *sp = lv.species.data() + size_t(k) * stride;
      if (!systematic) {
        auto &bucket = seen[hash_bytes(sp, stride)];
        bool dup = false;
        for (uint32_t o : bucket)
          dup |= std::memcmp(lv.species.data() + size_t(o) * stride, sp, stride) == 0;
        if (dup) continue;
        bucket.push_back(k);
      }
      R.cand.push_back(lv.cand[k]);
      R.sub.insert(R.sub.end(), lv.sub.begin() + size_t(k) * nprob,
                   lv.sub.begin() + size_t(k + 1) * nprob);
      R.species.insert(R.species.end(), sp, sp + stride);
      ++kept;
    }
    R.level_obj.push_back(obj);
    R.level_start.push_back(R.cand.size());
  }
  job->levels.clear();
  // SRO cubes are re-scored on the device: eagerly for packs of ordinary size
  if (R.size() && R.size() <= (uint64_t(1) << 20)) {
    std::vector<double> sro(R.size() * job->cube_total);
    result_sro(job, 0, R.size(), sro.data());
    R.sro = std::move(sro);
  }
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

void result_sro(sqsg_job *job, uint64_t first, uint64_t count, double *out) {
  const ResultSet &R = job->res;
  if (first + count > R.size()) throw std::out_of_range("result index out of range");
  if (!R.sro.empty()) {
    std::memcpy(out, R.sro.data() + first * job->cube_total,
                count * job->cube_total * sizeof(double));
    return;
  }
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
                  d_obj.as<double>(), d_sro.as<double>(), SQSG_KERNEL_SPARSE_WAVE, job->stream);
      std::vector<double> obj(B), sro(B * nb);
      SQSG_HIP(hipMemcpyAsync(obj.data(), d_obj.as<void>(), B * sizeof(double),
                              hipMemcpyDeviceToHost, job->stream));
      SQSG_HIP(hipMemcpyAsync(sro.data(), d_sro.as<void>(), sro.size() * sizeof(double),
                              hipMemcpyDeviceToHost, job->stream));
      SQSG_HIP(hipStreamSynchronize(job->stream));
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
