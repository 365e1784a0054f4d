// Sparse pair-list scoring kernels for gfx950 (wave64).
//
//  * score_explicit_kernel: explicit candidates -> ordered int32 bond histogram,
//    FP64 objective and SRO cube (core/optimization.h:74-103).
//  * run_wave_kernel: the plain single-job kernel.  Each wavefront generates a group
//    of candidates lane-parallel into LDS rows (MC: counter-based rapidrand
//    Fisher-Yates, core/shuffle.h:23-50; systematic: 128-bit unrank then lexicographic
//    successor, core/permutation.h:52-102), then scores them one at a time with all 64
//    lanes striding the pair list into an LDS histogram, computes the FP64 objective in
//    the reference's term order and appends candidates at or below the tie-keeping
//    threshold to a record sink.
//
// Compiled with -ffp-contract=off: objectives must be bit-identical to the reference's
// sequential, non-fused evaluation because results are keyed on exact objective values.
#include <hip/hip_runtime.h>

#include "candidates.hpp"
#include "kernels.hpp"

namespace sqsg {
namespace {

// copy the sorted configuration (4-byte padded) into an LDS row
__device__ __forceinline__ void load_row(uint8_t *row, const DevProblem &pr) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(pr.species0);
  uint32_t *dst = reinterpret_cast<uint32_t *>(row);
  const uint32_t words = (pr.N + 3) >> 2;
  for (uint32_t w = 0; w < words; ++w) dst[w] = src[w];
}

// count_bonds (optimization.h:74-77) over the pair list into the wave's LDS histogram
__device__ __forceinline__ void count_wave(const uint8_t *row, const DevProblem &pr,
                                           uint32_t *hist, int lane) {
  const uint32_t K = pr.K, nb = pr.S * K * K;
  for (uint32_t b = lane; b < nb; b += 64) hist[b] = 0;
  wave_sync();
  const uint32_t *pairs = pr.pairs;
  const uint32_t P = pr.P;
  uint32_t p = lane;
  for (; p + 192 < P; p += 256) {
    const uint32_t q0 = pairs[p], q1 = pairs[p + 64], q2 = pairs[p + 128], q3 = pairs[p + 192];
    const uint32_t a0 = row[q0 & 0xfff], b0 = row[(q0 >> 12) & 0xfff];
    const uint32_t a1 = row[q1 & 0xfff], b1 = row[(q1 >> 12) & 0xfff];
    const uint32_t a2 = row[q2 & 0xfff], b2 = row[(q2 >> 12) & 0xfff];
    const uint32_t a3 = row[q3 & 0xfff], b3 = row[(q3 >> 12) & 0xfff];
    atomicAdd(&hist[((q0 >> 24) * K + a0) * K + b0], 1u);
    atomicAdd(&hist[((q1 >> 24) * K + a1) * K + b1], 1u);
    atomicAdd(&hist[((q2 >> 24) * K + a2) * K + b2], 1u);
    atomicAdd(&hist[((q3 >> 24) * K + a3) * K + b3], 1u);
  }
  for (; p < P; p += 64) {
    const uint32_t q = pairs[p];
    atomicAdd(&hist[((q >> 24) * K + row[q & 0xfff]) * K + row[(q >> 12) & 0xfff]], 1u);
  }
  wave_sync();
}

// compute_objective (optimization.h:79-103): lanes evaluate the terms, lane 0 sums them in
// the reference order (s, xi, then the diagonal and eta > xi), without contraction
__device__ __forceinline__ double objective_wave(const DevProblem &pr, const uint32_t *hist,
                                                 double *terms, double *sro, int lane) {
  const uint32_t K = pr.K;
  for (uint32_t t = lane; t < pr.T; t += 64) {
    const uint32_t d = pr.terms[t];
    const uint32_t s = d & 0xff, xi = (d >> 8) & 0xff, eta = d >> 16;
    const uint32_t ix = (s * K + xi) * K + eta;
    uint32_t cnt = hist[ix];
    if (xi != eta) cnt += hist[(s * K + eta) * K + xi];
    double v;
    terms[t] = term_value(pr.single != 0, cnt, pr.pref[ix], pr.pw[ix], pr.tgt[ix], v);
    if (sro) {
      sro[ix] = v;
      if (xi != eta) sro[(s * K + eta) * K + xi] = v;
    }
  }
  wave_sync();
  double obj = 0.0;
  if (lane == 0) {
    if (pr.single) {  // configuration<float>: the sum in float
      float o = 0.0f;
      for (uint32_t t = 0; t < pr.T; ++t) o += static_cast<float>(terms[t]);
      obj = o;
    } else {
      for (uint32_t t = 0; t < pr.T; ++t) obj += terms[t];
    }
  }
  obj = __shfl(obj, 0);
  return obj;
}

__global__ __launch_bounds__(256) void score_explicit_kernel(ScoreParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  uint8_t *wl = lds + wave * p.wave_lds;
  uint8_t *row = wl + p.off_row;
  uint32_t *hist = reinterpret_cast<uint32_t *>(wl + p.off_hist);
  double *terms = reinterpret_cast<double *>(wl + p.off_terms);
  const DevProblem &pr = p.prob;
  const uint32_t N = pr.N, nb = pr.S * pr.K * pr.K;
  for (uint64_t b = static_cast<uint64_t>(blockIdx.x) * W + wave; b < p.B;
       b += static_cast<uint64_t>(gridDim.x) * W) {
    for (uint32_t i = lane; i < N; i += 64) row[i] = p.species[b * N + i];
    wave_sync();
    count_wave(row, pr, hist, lane);
    if (p.bonds)
      for (uint32_t k = lane; k < nb; k += 64) p.bonds[b * nb + k] = static_cast<int32_t>(hist[k]);
    const double obj = objective_wave(pr, hist, terms, p.sro ? p.sro + b * nb : nullptr, lane);
    if (lane == 0) p.objective[b] = obj;
    wave_sync();
  }
}

__global__ __launch_bounds__(256) void run_wave_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6, W = blockDim.x >> 6;
  uint8_t *wl = lds + wave * p.wave_lds;
  uint32_t *hist = reinterpret_cast<uint32_t *>(wl + p.off_hist);
  double *terms = reinterpret_cast<double *>(wl + p.off_terms);
  const uint64_t gw = static_cast<uint64_t>(blockIdx.x) * W + wave;
  const uint64_t GW = static_cast<uint64_t>(gridDim.x) * W;
  const uint32_t G = p.gen_group;
  const uint64_t per_seg = static_cast<uint64_t>(G) * p.rounds;
  const uint64_t total = p.cand_end - p.cand_begin;
  const uint64_t nseg = (total + per_seg - 1) / per_seg;
  double lvl[kMaxLevels];
  uint32_t nlvl = 0;
  unsigned long long scored = 0;
  for (uint64_t seg = gw; seg < nseg; seg += GW) {
    const uint64_t sbeg = p.cand_begin + seg * per_seg;
    for (uint32_t r = 0; r < p.rounds; ++r) {
      if (static_cast<uint32_t>(lane) < G) {
        const uint64_t cand = p.systematic ? sbeg + static_cast<uint64_t>(lane) * p.rounds + r
                                           : sbeg + static_cast<uint64_t>(r) * G + lane;
        if (cand < p.cand_end) {
          for (uint32_t q = 0; q < p.nprob; ++q) {
            const DevProblem &pr = p.prob[q];
            uint8_t *row = wl + p.off_rows[q] + lane * pr.row_stride;
            if (p.systematic) {
              if (r == 0) {
                load_row(row, pr);
                unrank_candidate(Strided<0>{row}, pr, cand + 1, 1u);
              } else {
                next_perm_candidate(Strided<0>{row}, pr.bounds[0], pr.bounds[1] - pr.bounds[0]);
              }
            } else {
              load_row(row, pr);
              if (!(p.chunk_quirk && block_start(p, cand)))
                shuffle_candidate(Strided<0>{row}, pr, cand);
            }
          }
        }
      }
      wave_sync();
      for (uint32_t c = 0; c < G; ++c) {
        const uint64_t cand = p.systematic ? sbeg + static_cast<uint64_t>(c) * p.rounds + r
                                           : sbeg + static_cast<uint64_t>(r) * G + c;
        if (cand >= p.cand_end) continue;
        double sub[kMaxProblems];
        for (uint32_t q = 0; q < p.nprob; ++q) {
          const DevProblem &pr = p.prob[q];
          const uint8_t *row = wl + p.off_rows[q] + c * pr.row_stride;
          count_wave(row, pr, hist, lane);
          sub[q] = objective_wave(pr, hist, terms, nullptr, lane);
          wave_sync();
        }
        double tot;
        if (p.nprob == 1) {
          tot = sub[0];
        } else {  // helpers::sum folds from T(0) (core/helpers/fold.h:51-53)
          tot = 0.0;
          for (uint32_t q = 0; q < p.nprob; ++q) tot = add_term(p.prob[0].single != 0, tot, sub[q]);
        }
        ++scored;
        const unsigned long long key = okey(tot);
        const unsigned long long g =
            p.thr_key;
        // the wave's L smallest distinct objectives of its earlier candidates: a candidate
        // above the L-th has L smaller distinct objectives before it and is not inserted
        const unsigned long long lk =
            p.wave_filter && nlvl == p.L ? okey(lvl[p.L - 1]) : ~0ull;
        if (p.wave_filter && key <= g) {
          bool present = false;
          for (uint32_t i = 0; i < nlvl; ++i) present |= (lvl[i] == tot);
          if (!present && (nlvl < p.L || tot < lvl[nlvl - 1])) {
            uint32_t pos = nlvl < p.L ? nlvl++ : nlvl - 1;
            while (pos > 0 && lvl[pos - 1] > tot) {
              lvl[pos] = lvl[pos - 1];
              --pos;
            }
            lvl[pos] = tot;
          }
        }
        if (key <= g && key <= lk) {
          uint32_t slot = 0;
          if (lane == 0) slot = atomicAdd(p.rec_count, 1u);
          slot = __shfl(slot, 0);
          if (slot < p.rec_cap) {
            uint8_t *dst = p.rec_species + static_cast<size_t>(slot) * p.rec_stride;
            uint32_t off = 0;
            for (uint32_t q = 0; q < p.nprob; ++q) {
              const DevProblem &pr = p.prob[q];
              const uint8_t *row = wl + p.off_rows[q] + c * pr.row_stride;
              for (uint32_t i = lane; i < pr.N; i += 64) dst[off + i] = row[i];
              off += pr.N;
            }
            if (lane == 0) {
              p.rec_cand[slot] = cand;
              p.rec_obj[slot] = tot;
              for (uint32_t q = 0; q < p.nprob; ++q) p.rec_sub[slot * p.nprob + q] = sub[q];
            }
          }
        }
      }
      wave_sync();
    }
  }
  if (lane == 0 && scored) atomicAdd(p.scored, scored);
}

hipError_t set_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

}  // namespace

hipError_t launch_score_explicit(const ScoreParams &p, hipStream_t stream) {
  if (p.B == 0) return hipSuccess;
  const uint32_t W = 4;
  const size_t lds = static_cast<size_t>(W) * p.wave_lds;
  hipError_t e = set_lds(reinterpret_cast<const void *>(&score_explicit_kernel), lds);
  if (e != hipSuccess) return e;
  uint32_t blocks = (p.B + W - 1) / W;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(score_explicit_kernel, dim3(blocks), dim3(64 * W), lds, stream, p);
  return hipGetLastError();
}

hipError_t launch_run_wave(const RunParams &p, uint32_t blocks, uint32_t waves_per_block,
                           hipStream_t stream) {
  const size_t lds = static_cast<size_t>(waves_per_block) * p.wave_lds;
  hipError_t e = set_lds(reinterpret_cast<const void *>(&run_wave_kernel), lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(run_wave_kernel, dim3(blocks), dim3(64 * waves_per_block), lds, stream, p);
  return hipGetLastError();
}

}  // namespace sqsg
