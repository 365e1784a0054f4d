// Lane-per-candidate sparse scoring for gfx950: each of the 64 lanes of a wavefront owns
// one candidate; the wavefront walks the neighbour structure uniformly (scalar loads), so
// the pair list costs no per-candidate memory traffic.
//
// LDS of one wavefront:
//   rows  [site][64 lanes] bytes      species * F of every lane's candidate
//   hist  [a * F + w][64 lanes] u32   16-bit count fields of the current shell
// For every site i and shell s the lane counts its upper neighbours j > i by species in
// F-bit fields of one register (acc += 1 << (F * b_j)), then adds the fields to row
// a = species(i) of the shell histogram.  A field of row a, column b therefore ends up as
// bonds(s, a, b) of core/optimization.h:74-77 (ordered, a = species[i], i < j).  Shells are
// processed in order and each shell's objective terms are added as soon as its histogram
// is complete, which is exactly the summation order of optimization.h:85-100.
#include <hip/hip_runtime.h>

#include "candidates.hpp"
#include "kernels.hpp"

namespace sqsg {
namespace {

// F: bits per neighbour-count field.  K * F <= 30 keeps bits 30..31 free for the padding
// entries of a group (at most 3, they point to a dummy site whose value is K * F).  The
// histogram keeps W = ceil(K / 2) words of two 16-bit fields per species row; the row
// address is derived from the site value a * F, so W * 256 must be a multiple of F.
constexpr uint32_t lane_bits(int km) {
  const uint32_t w = (km + 1) / 2;
  uint32_t f = 30u / static_cast<uint32_t>(km);
  if (f > 8) f = 8;
  while (f > 1 && (w * 256u) % f) --f;
  return f;
}

template <int KM> struct LaneShape {
  static constexpr uint32_t F = lane_bits(KM);
  static constexpr uint32_t W = (KM + 1) / 2;
  static constexpr uint32_t MASK = (1u << F) - 1u;
  static constexpr uint32_t ROW_BYTES_PER_UNIT = W * 256u / F;  // bytes per unit of a * F
};

__device__ __forceinline__ void load_col(uint8_t *col, const DevProblem &pr) {
  for (uint32_t i = 0; i < pr.N; ++i) col[i << 6] = pr.species0F[i];
}

template <int KM>
__device__ __forceinline__ uint32_t field_count(const uint32_t *hl, uint32_t a, uint32_t b) {
  constexpr uint32_t W = LaneShape<KM>::W;
  return (hl[(a * W + (b >> 1)) << 6] >> ((b & 1u) << 4)) & 0xffffu;
}

// Bond histogram + objective of the lane's candidate (`col`), shell by shell.  `hl` is
// the lane's histogram column (hist + lane).  Optionally writes ordered bonds and SRO.
// SINGLE: the terms and their sum in float (configuration<float>), fixed at compile time so
// that the double path carries no per-term selection.
template <int KM, bool SINGLE>
__device__ double score_lane_t(const uint8_t *col, const DevProblem &pr, uint32_t *hl,
                               int32_t *bonds, double *sro) {
  using Sh = LaneShape<KM>;
  constexpr uint32_t F = Sh::F, W = Sh::W, MASK = Sh::MASK;
  const uint32_t K = pr.K;
  // the neighbour stream is read with in-order vector loads (bounds-checked: reads past
  // the end return zeros), so loads of the next block overlap the LDS gathers of this one
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t *>(pr.lane_stream), static_cast<short>(0),
      static_cast<int>(pr.lane_off[pr.S] * 16u), 0x00020000);
  uint8_t *hb = reinterpret_cast<uint8_t *>(hl);
  double obj = 0.0;
  for (uint32_t s = 0; s < pr.S; ++s) {
#pragma unroll
    for (uint32_t a = 0; a < KM; ++a)
#pragma unroll
      for (uint32_t w = 0; w < W; ++w)
        if (a < K) hl[(a * W + w) << 6] = 0u;
    uint32_t blk = pr.lane_off[s];
    const uint32_t end = pr.lane_off[s + 1];
    auto hdr = __builtin_amdgcn_raw_buffer_load_b128(rs, blk * 16u, 0, 0);
    while (blk < end) {
      const uint32_t i_off = hdr[0], nb = hdr[1];
      ++blk;
      const uint32_t a6 = col[i_off];
      uint32_t acc = 0;
      auto d0 = __builtin_amdgcn_raw_buffer_load_b128(rs, blk * 16u, 0, 0);
      auto d1 = __builtin_amdgcn_raw_buffer_load_b128(rs, (blk + 1) * 16u, 0, 0);
      for (uint32_t t = 0; t < nb; ++t) {
        const auto d2 = __builtin_amdgcn_raw_buffer_load_b128(rs, (blk + 2) * 16u, 0, 0);
        const uint32_t b0 = col[d0[0]], b1 = col[d0[1]], b2 = col[d0[2]], b3 = col[d0[3]];
        acc += ((1u << b0) + (1u << b1)) + ((1u << b2) + (1u << b3));
        d0 = d1;
        d1 = d2;
        ++blk;
      }
      hdr = d0;  // the block after a group's data is the next group's header
      uint8_t *h = hb + a6 * Sh::ROW_BYTES_PER_UNIT;
#pragma unroll
      for (uint32_t w = 0; w < W; ++w) {
        uint32_t v = (acc >> (2 * w * F)) & MASK;
        if (2 * w + 1 < KM) v |= ((acc >> ((2 * w + 1) * F)) & MASK) << 16;
        atomicAdd(reinterpret_cast<uint32_t *>(h + (w << 8)), v);
      }
    }
    // this shell's objective terms, optimization.h:85-100
    constexpr bool single = SINGLE;
    const double *pref = pr.pref + s * K * K;
    const double *pw = pr.pw + s * K * K;
    const double *tgt = pr.tgt + s * K * K;
    for (uint32_t xi = 0; xi < K; ++xi) {
      const uint32_t cxx = field_count<KM>(hl, xi, xi);
      double sxx;
      obj = add_term(single, obj,
                     term_value(single, cxx, pref[xi * K + xi], pw[xi * K + xi], tgt[xi * K + xi], sxx));
      if (sro) sro[(s * K + xi) * K + xi] = sxx;
      if (bonds) bonds[(s * K + xi) * K + xi] = static_cast<int32_t>(cxx);
      for (uint32_t eta = xi + 1; eta < K; ++eta) {
        const uint32_t cab = field_count<KM>(hl, xi, eta), cba = field_count<KM>(hl, eta, xi);
        double v;
        obj = add_term(single, obj, term_value(single, cab + cba, pref[xi * K + eta],
                                               pw[xi * K + eta], tgt[xi * K + eta], v));
        if (sro) sro[(s * K + xi) * K + eta] = v, sro[(s * K + eta) * K + xi] = v;
        if (bonds) {
          bonds[(s * K + xi) * K + eta] = static_cast<int32_t>(cab);
          bonds[(s * K + eta) * K + xi] = static_cast<int32_t>(cba);
        }
      }
    }
  }
  return obj;
}

template <int KM>
__device__ __forceinline__ double score_lane(const uint8_t *col, const DevProblem &pr,
                                             uint32_t *hl, int32_t *bonds, double *sro) {
  return pr.single ? score_lane_t<KM, true>(col, pr, hl, bonds, sro)
                   : score_lane_t<KM, false>(col, pr, hl, bonds, sro);
}

__device__ __forceinline__ void init_rows(const RunParams &p, uint8_t *wl, int lane) {
  // dummy site N of every problem: padding entries of the neighbour stream read it
  for (uint32_t q = 0; q < p.nprob; ++q)
    wl[p.off_rows[q] + (p.prob[q].N << 6) + lane] =
        static_cast<uint8_t>(p.prob[q].K * p.prob[q].F);
}

// One segment of a job: 64 lanes x p.rounds candidates starting at sbeg, candidates at or
// beyond `end` are inactive.  Generates, scores, and appends candidates at or below the
// job's threshold to its record sink.
template <int KM>
__device__ void lane_segment(const RunParams &p, uint64_t sbeg, uint64_t end, uint8_t *wl,
                             uint32_t *hl, int lane, LevelFilter &lf,
                             unsigned long long &scored) {
  for (uint32_t r = 0; r < p.rounds; ++r) {
    const uint64_t cand = p.systematic ? sbeg + static_cast<uint64_t>(lane) * p.rounds + r
                                       : sbeg + static_cast<uint64_t>(r) * 64 + lane;
    const bool active = cand < end;
    // every lane keeps a valid configuration in its column (inactive lanes score the
    // sorted one and discard it), so every LDS address below stays in the wave's region
    for (uint32_t q = 0; q < p.nprob; ++q) {
      const DevProblem &pr = p.prob[q];
      uint8_t *col = wl + p.off_rows[q] + lane;
      if (p.systematic) {
        if (r == 0) {
          load_col(col, pr);
          if (active) unrank_candidate(Strided<6>{col}, pr, cand + 1, pr.F);
        } else if (active) {
          next_perm_candidate(Strided<6>{col}, pr.bounds[0], pr.bounds[1] - pr.bounds[0]);
        }
      } else {
        load_col(col, pr);
        if (active && !(p.chunk_quirk && block_start(p, cand)))
          shuffle_candidate(Strided<6>{col}, pr, cand);
      }
    }
    if (__ballot(active) == 0) break;  // the tail of a systematic segment
    double sub[kMaxProblems];
    for (uint32_t q = 0; q < p.nprob; ++q)
      sub[q] = score_lane<KM>(wl + p.off_rows[q] + lane, p.prob[q], hl, nullptr, nullptr);
    double tot;
    if (p.nprob == 1) {
      tot = sub[0];
    } else {  // helpers::sum folds from T(0) (core/helpers/fold.h:51-53)
      tot = 0.0;
      for (uint32_t q = 0; q < p.nprob; ++q) tot = add_term(p.prob[0].single != 0, tot, sub[q]);
    }
    scored += static_cast<unsigned long long>(__popcll(__ballot(active)));
    const unsigned long long g =
        p.thr_key;
    // a candidate above the threshold, or above the filter's L-th level when its turn
    // comes, cannot be inserted
    const bool below = active && okey(tot) <= g;
    const bool pass = (filter_pass(lf, p, __ballot(below), tot) >> lane) & 1ull;
    uint32_t slot = ~0u;
    if (pass) {
      slot = atomicAdd(p.rec_count, 1u);
      if (slot < p.rec_cap) {
        p.rec_cand[slot] = cand;
        p.rec_obj[slot] = tot;
        for (uint32_t q = 0; q < p.nprob; ++q) p.rec_sub[slot * p.nprob + q] = sub[q];
      }
    }
    // species of the recorded candidates, one column (candidate) at a time by the wavefront
    for (unsigned long long m = __ballot(pass); m; m &= m - 1) {
      const int l = __ffsll(m) - 1;
      const uint32_t s = __shfl(slot, l);
      if (s >= p.rec_cap) continue;
      uint8_t *dst = p.rec_species + static_cast<size_t>(s) * p.rec_stride;
      uint32_t off = 0;
      for (uint32_t q = 0; q < p.nprob; ++q) {
        const DevProblem &pr = p.prob[q];
        const uint8_t *col = wl + p.off_rows[q] + l;
        for (uint32_t i = static_cast<uint32_t>(lane); i < pr.N; i += 64)
          dst[off + i] = col[i << 6] / pr.F;
        off += pr.N;
      }
    }
  }
}

template <int KM>
__global__ __launch_bounds__(256) void run_lane_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6, NW = blockDim.x >> 6;
  uint8_t *wl = lds + wave * p.wave_lds;
  uint32_t *hl = reinterpret_cast<uint32_t *>(wl + p.off_hist) + lane;
  const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * NW + wave;
  const uint64_t GW = static_cast<uint64_t>(gridDim.x) * NW;
  const uint64_t per_seg = 64ull * p.rounds;
  const uint64_t total = p.cand_end - p.cand_begin;
  const uint64_t nseg = (total + per_seg - 1) / per_seg;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  init_rows(p, wl, lane);
  for (uint64_t seg = gw; seg < nseg; seg += GW)
    lane_segment<KM>(p, p.cand_begin + seg * per_seg, p.cand_end, wl, hl, lane, lf, scored);
  if (lane == 0 && scored) atomicAdd(p.scored, scored);
}

// Persistent form: every wavefront claims one-segment chunks of the jobs' candidate ranges
// until the work list is drained; jobs of any size, species and shell count (up to KM
// species) share the launch.  Each job has its own threshold and record sink.
template <int KM>
__global__ __launch_bounds__(256) void batch_lane_kernel(BatchParams bp) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6;
  uint8_t *wl = lds + wave * bp.wave_lds;
  const uint32_t num_items = *bp.num_items;
  uint32_t last_job = 0xffffffffu, it = 0, job = 0;
  uint64_t cbeg = 0, cend = 0;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  while (claim_chunk(bp, num_items, it, lane, job, cbeg, cend)) {
    const RunParams &p = bp.jobs[job];
    if (job != last_job) {
      if (last_job != 0xffffffffu && lane == 0 && scored)
        atomicAdd(bp.jobs[last_job].scored, scored);
      scored = 0;
      lf.n = 0;
      last_job = job;
      init_rows(p, wl, lane);
    }
    uint32_t *hl = reinterpret_cast<uint32_t *>(wl + p.off_hist) + lane;
    lane_segment<KM>(p, cbeg, cend, wl, hl, lane, lf, scored);
  }
  if (last_job != 0xffffffffu && lane == 0 && scored) atomicAdd(bp.jobs[last_job].scored, scored);
}

// After a persistent step: copies each job's record and scored counts and its first
// stage_cap records into the job's stage in host memory (one block per job, coalesced
// writes), so the host reads them without a copy that would wait for the next step's
// kernel.
__global__ __launch_bounds__(256) void batch_stage_kernel(BatchParams bp) {
  const uint32_t t = blockIdx.x;
  if (t >= *bp.num_jobs) return;
  const RunParams &p = bp.jobs[t];
  if (!p.stage) return;
  const unsigned int count = *p.rec_count;
  // the host reads the records from the stage only when all of them fit it; more (a job's
  // first steps) are fetched from the sink
  const uint32_t n = count <= min(p.rec_cap, p.stage_cap) ? count : 0u;
  uint8_t *st = p.stage;
  if (threadIdx.x == 0) {
    *reinterpret_cast<unsigned int *>(st) = count;
    *reinterpret_cast<unsigned long long *>(st + 8) = *p.scored;
  }
  uint64_t *cand = reinterpret_cast<uint64_t *>(st + 16);
  double *obj = reinterpret_cast<double *>(cand + p.stage_cap);
  double *sub = obj + p.stage_cap;
  uint8_t *sp = reinterpret_cast<uint8_t *>(sub + static_cast<size_t>(p.stage_cap) * p.nprob);
  for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
    cand[k] = p.rec_cand[k];
    obj[k] = p.rec_obj[k];
  }
  for (uint32_t k = threadIdx.x; k < n * p.nprob; k += blockDim.x) sub[k] = p.rec_sub[k];
  // species in 8-byte words (both regions start 8-byte aligned), then the tail bytes
  const uint32_t bytes = n * p.rec_stride, words = bytes / 8;
  for (uint32_t k = threadIdx.x; k < words; k += blockDim.x)
    reinterpret_cast<uint64_t *>(sp)[k] = reinterpret_cast<const uint64_t *>(p.rec_species)[k];
  for (uint32_t k = 8 * words + threadIdx.x; k < bytes; k += blockDim.x) sp[k] = p.rec_species[k];
}

// zeroes the item claim counters and every job's record and scored counters (a kernel, so
// the reset can sit inside a captured graph)
__global__ void batch_reset_kernel(BatchParams bp, uint32_t max_jobs) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < bp.max_items && t < *bp.num_items) bp.claims[t] = 0u;
  if (t < max_jobs && t < *bp.num_jobs) {
    *bp.jobs[t].rec_count = 0u;
    *bp.jobs[t].scored = 0ull;
  }
}

template <int KM>
__global__ __launch_bounds__(256) void score_lane_kernel(ScoreParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6;
  uint8_t *wl = lds + wave * p.wave_lds;
  uint8_t *col = wl + p.off_row + lane;
  uint32_t *hl = reinterpret_cast<uint32_t *>(wl + p.off_hist) + lane;
  const DevProblem &pr = p.prob;
  const uint32_t nb = pr.S * pr.K * pr.K;
  const uint64_t b = (static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + wave) * 64 + lane;
  const bool active = b < p.B;
  for (uint32_t i = 0; i < pr.N; ++i)
    col[i << 6] = active ? static_cast<uint8_t>(p.species[b * pr.N + i] * pr.F) : 0;
  col[pr.N << 6] = static_cast<uint8_t>(pr.K * pr.F);
  const double obj = score_lane<KM>(col, pr, hl, active && p.bonds ? p.bonds + b * nb : nullptr,
                                    active && p.sro ? p.sro + b * nb : nullptr);
  if (active) p.objective[b] = obj;
}

template <int KM> hipError_t launch_run_lane_k(const RunParams &p, uint32_t blocks, uint32_t W,
                                               hipStream_t s) {
  const size_t lds = static_cast<size_t>(W) * p.wave_lds;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&run_lane_kernel<KM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(run_lane_kernel<KM>, dim3(blocks), dim3(64 * W), lds, s, p);
  return hipGetLastError();
}

template <int KM> hipError_t launch_score_lane_k(const ScoreParams &p, hipStream_t s) {
  const uint32_t W = 1;
  const size_t lds = p.wave_lds;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&score_lane_kernel<KM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  const uint32_t blocks = (p.B + 63) / 64;
  hipLaunchKernelGGL(score_lane_kernel<KM>, dim3(blocks), dim3(64 * W), lds, s, p);
  return hipGetLastError();
}

template <int KM> hipError_t launch_batch_lane_k(const BatchParams &bp, uint32_t blocks,
                                                 uint32_t W, hipStream_t s) {
  const size_t lds = static_cast<size_t>(W) * bp.wave_lds;
  if (lds > 65536) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_lane_kernel<KM>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(lds));
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(batch_lane_kernel<KM>, dim3(blocks), dim3(64 * W), lds, s, bp);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_stage(const BatchParams &bp, uint32_t max_jobs, hipStream_t s) {
  hipLaunchKernelGGL(batch_stage_kernel, dim3(max_jobs), dim3(256), 0, s, bp);
  return hipGetLastError();
}

hipError_t launch_batch_reset(const BatchParams &bp, uint32_t max_jobs, hipStream_t s) {
  const uint32_t n = max_jobs > bp.max_items ? max_jobs : bp.max_items;
  const uint32_t threads = 256, blocks = (n + threads - 1) / threads;
  hipLaunchKernelGGL(batch_reset_kernel, dim3(blocks), dim3(threads), 0, s, bp, max_jobs);
  return hipGetLastError();
}

hipError_t prepare_batch_lane(uint32_t kmax, size_t lds) {
  if (lds <= 65536) return hipSuccess;
  const void *fn = nullptr;
  switch (kmax) {
    case 1: case 2: fn = reinterpret_cast<const void *>(&batch_lane_kernel<2>); break;
    case 3: fn = reinterpret_cast<const void *>(&batch_lane_kernel<3>); break;
    case 4: fn = reinterpret_cast<const void *>(&batch_lane_kernel<4>); break;
    case 5: fn = reinterpret_cast<const void *>(&batch_lane_kernel<5>); break;
    case 6: fn = reinterpret_cast<const void *>(&batch_lane_kernel<6>); break;
    case 7: fn = reinterpret_cast<const void *>(&batch_lane_kernel<7>); break;
    case 8: fn = reinterpret_cast<const void *>(&batch_lane_kernel<8>); break;
    default: return hipErrorInvalidValue;
  }
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(lds));
}

hipError_t launch_batch_lane(const BatchParams &bp, uint32_t kmax, uint32_t blocks,
                             uint32_t waves_per_block, hipStream_t s) {
  switch (kmax) {
    case 1: case 2: return launch_batch_lane_k<2>(bp, blocks, waves_per_block, s);
    case 3: return launch_batch_lane_k<3>(bp, blocks, waves_per_block, s);
    case 4: return launch_batch_lane_k<4>(bp, blocks, waves_per_block, s);
    case 5: return launch_batch_lane_k<5>(bp, blocks, waves_per_block, s);
    case 6: return launch_batch_lane_k<6>(bp, blocks, waves_per_block, s);
    case 7: return launch_batch_lane_k<7>(bp, blocks, waves_per_block, s);
    case 8: return launch_batch_lane_k<8>(bp, blocks, waves_per_block, s);
    default: return hipErrorInvalidValue;
  }
}

uint32_t lane_field_bits(uint32_t kmax) { return lane_bits(kmax < 2 ? 2 : static_cast<int>(kmax)); }

hipError_t launch_run_lane(const RunParams &p, uint32_t kmax, uint32_t blocks,
                           uint32_t waves_per_block, hipStream_t s) {
  switch (kmax) {
    case 1: case 2: return launch_run_lane_k<2>(p, blocks, waves_per_block, s);
    case 3: return launch_run_lane_k<3>(p, blocks, waves_per_block, s);
    case 4: return launch_run_lane_k<4>(p, blocks, waves_per_block, s);
    case 5: return launch_run_lane_k<5>(p, blocks, waves_per_block, s);
    case 6: return launch_run_lane_k<6>(p, blocks, waves_per_block, s);
    case 7: return launch_run_lane_k<7>(p, blocks, waves_per_block, s);
    case 8: return launch_run_lane_k<8>(p, blocks, waves_per_block, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_score_lane(const ScoreParams &p, uint32_t kmax, hipStream_t s) {
  if (p.B == 0) return hipSuccess;
  switch (kmax) {
    case 1: case 2: return launch_score_lane_k<2>(p, s);
    case 3: return launch_score_lane_k<3>(p, s);
    case 4: return launch_score_lane_k<4>(p, s);
    case 5: return launch_score_lane_k<5>(p, s);
    case 6: return launch_score_lane_k<6>(p, s);
    case 7: return launch_score_lane_k<7>(p, s);
    case 8: return launch_score_lane_k<8>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sqsg
