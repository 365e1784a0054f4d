// Candidate generation and tie-keeping helpers shared by the scoring kernels.
//
// A candidate's species live in LDS as N bytes with a fixed stride: contiguous rows
// (stride 1: the wave and dense kernels) or one column of a [site][64 lanes] block
// (stride 64: the lane kernel).  The generators below are the device definition of a
// candidate, identical for every kernel:
//  * MC candidate t: Fisher-Yates of the sorted configuration with draws
//    [t * draws, (t + 1) * draws) of the rapidrand stream (core/shuffle.h:23-50; the stream
//    is a counter, state_k = seed + k * s0), windows in bounds order, i from w down to 2;
//  * systematic rank r: the 1-based lexicographic unrank of the window bounds[0..1]
//    (core/permutation.h:52-78) and its successors (core/permutation.h:80-102).
#pragma once

#include "device_common.hpp"
#include "device_params.hpp"

namespace sqsg {
namespace {

template <uint32_t SH> struct Strided {
  uint8_t *p;
  __device__ __forceinline__ uint8_t &operator[](uint32_t i) const { return p[i << SH]; }
};

// PR is DevProblem, possibly qualified with the constant address space (a job table entry
// read with scalar loads)
// (r * n) >> 32 over 64 bits is the high word of the 32 x 32 product, one v_mul_hi_u32.
// Written in C the product stays 64-bit in the compiler's view, and loop strength
// reduction then carries the shuffle's counter as a 64-bit value with an extra
// quarter-rate v_mad_u64_u32 per step; the one-instruction asm (a plain VALU operation,
// no memory access) keeps it a 32-bit multiply.
__device__ __forceinline__ uint32_t bounded(uint32_t r, uint32_t n) {  // shuffle.h:28-32
  uint32_t q;
  asm("v_mul_hi_u32 %0, %1, %2" : "=v"(q) : "v"(r), "v"(n));
  return q;
}

// The draw of step i - 1 is computed between the reads and the writes of step i, so its
// multiplies cover the LDS round trip of the swap (the byte reads and writes of one lane
// stay in program order, and the draws are consumed in the same order as a plain loop).
template <uint32_t SH, class PR>
__device__ __forceinline__ void shuffle_candidate(Strided<SH> v, const PR &pr, uint64_t cand) {
  uint64_t st = pr.seed + (cand * static_cast<uint64_t>(pr.draws)) * kSecret0;
  for (uint32_t b = 0; b < pr.nbounds; ++b) {
    const uint32_t lo = pr.bounds[2 * b], hi = pr.bounds[2 * b + 1];
    uint32_t i = hi - lo;
    if (i < 2) continue;
    uint32_t p = bounded(draw32(st), i);
    for (; i > 2; --i) {
      const uint8_t a = v[lo + i - 1], c = v[lo + p];
      const uint32_t pn = bounded(draw32(st), i - 1);
      v[lo + i - 1] = c;
      v[lo + p] = a;
      p = pn;
    }
    const uint8_t a = v[lo + 1], c = v[lo + p];  // i == 2
    v[lo + 1] = c;
    v[lo + p] = a;
  }
}

// writes the packed species of rank `rank` (1-based) times `scale` into the window
template <uint32_t SH, class PR>
__device__ void unrank_candidate(Strided<SH> v, const PR &pr, uint64_t rank, uint32_t scale) {
  uint32_t hist[kMaxLocalSpecies];
  for (uint32_t j = 0; j < kMaxLocalSpecies; ++j) hist[j] = pr.sys_hist[j];
  const uint32_t lo = pr.bounds[0], w = pr.bounds[1] - pr.bounds[0];
  uint64_t total = 1;
  uint32_t placed = 0;
  for (uint32_t j = 0; j < pr.sys_nspecies; ++j)
    for (uint32_t k = 1; k <= hist[j]; ++k) total = muldiv(total, ++placed, k);
  uint32_t k = 0;
  for (uint32_t i = 0; i < w; ++i)
    for (uint32_t j = 0; j < pr.sys_nspecies; ++j) {
      const uint64_t sc = muldiv(total, hist[j], w - i);
      if (rank <= sc) {
        v[lo + k++] = static_cast<uint8_t>(pr.sys_rev[j] * scale);
        total = sc;
        --hist[j];
        break;
      }
      rank -= sc;
    }
}

template <uint32_t SH>
__device__ void next_perm_candidate(Strided<SH> v, uint32_t lo, uint32_t w) {
  if (w < 2) return;
  int k = static_cast<int>(w) - 2;
  while (k >= 0 && v[lo + k] >= v[lo + k + 1]) --k;
  if (k < 0) return;
  int l = static_cast<int>(w) - 1;
  while (v[lo + k] >= v[lo + l]) --l;
  uint8_t t = v[lo + k];
  v[lo + k] = v[lo + l];
  v[lo + l] = t;
  for (int a = k + 1, b = static_cast<int>(w) - 1; a < b; ++a, --b) {
    t = v[lo + a];
    v[lo + a] = v[lo + b];
    v[lo + b] = t;
  }
}


// First candidate of a block of the reference's thread-pool split (sqs.h:461-466; the
// pool's blocks are near-equal with the first `remainder` ones one larger), which an
// interact Monte Carlo run scores unshuffled (sqs.h:374,400-402,439-440).
template <class RP>
__device__ __forceinline__ bool block_start(const RP &p, uint64_t cand) {
  return cand < p.blk_split ? cand % (p.blk_size + 1) == 0
                            : (cand - p.blk_split) % p.blk_size == 0;
}

// The N-byte LDS rows of the lanes in `pm` (row of lane l at rows + l * row_stride) into
// bytes [off, off + N) of their record slots (`slot` of each lane, rec_stride bytes per slot;
// slots at or past `cap` are dropped).  The wavefront copies one row at a time, 4 bytes per
// lane where the offsets allow (LDS rows are 16-byte aligned), else one byte per lane.
__device__ __forceinline__ void record_rows(uint8_t *rec_species, uint32_t rec_stride,
                                            uint32_t cap, unsigned long long pm, uint32_t slot,
                                            const uint8_t *rows, uint32_t row_stride,
                                            uint32_t N, uint32_t off, int lane) {
  const bool words = ((rec_stride | off | N | row_stride) & 3u) == 0;
  for (unsigned long long m = pm; m; m &= m - 1) {
    const int l = __ffsll(m) - 1;
    const uint32_t s = __shfl(slot, l);
    if (s >= cap) continue;
    uint8_t *dst = rec_species + static_cast<size_t>(s) * rec_stride + off;
    const uint8_t *row = rows + static_cast<uint32_t>(l) * row_stride;
    if (words) {
      for (uint32_t i = static_cast<uint32_t>(lane); i < N / 4; i += 64)
        reinterpret_cast<uint32_t *>(dst)[i] = reinterpret_cast<const uint32_t *>(row)[i];
    } else {
      for (uint32_t i = static_cast<uint32_t>(lane); i < N; i += 64) dst[i] = row[i];
    }
  }
}

// Bound of the wavefront filter as it stands before a round (it holds only candidates that
// precede the round's): the order key of its L-th level, or no bound.  The early exit uses it.
template <class RP>
__device__ __forceinline__ unsigned long long filter_bound(const LevelFilter &lf,
                                                           const RP &p) {
  return p.wave_filter && lf.n == p.L ? okey(lf.lvl[p.L - 1]) : ~0ull;
}

// Offers the objectives of the lanes in `mask` (a ballot; lanes hold ascending candidates)
// to the filter one lane at a time and returns the lanes that were at or below the filter's
// L-th level when their turn came (or met a filter of fewer than L levels): the others
// have L smaller distinct objectives among the candidates before them.  Without the filter,
// every lane of `mask` passes.
template <class RP>
__device__ __forceinline__ unsigned long long filter_pass(LevelFilter &lf, const RP &p,
                                                          unsigned long long mask,
                                                          double tot) {
  if (!p.wave_filter) return mask;
  // A lane above a full filter's L-th level neither passes nor changes the filter (that
  // level only decreases as lanes are offered), so only the others are walked: while the
  // job threshold is still open (a job's first steps) this skips nearly every lane.
  if (lf.n == p.L) mask &= __ballot(tot <= lf.lvl[p.L - 1]);
  unsigned long long pass = 0;
  while (mask) {
    const int l = __ffsll(mask) - 1;
    mask &= mask - 1;
    const double o = __shfl(tot, l);
    const bool full = lf.n == p.L;
    if (!full || o <= lf.lvl[p.L - 1]) pass |= 1ull << l;
    bool present = false;
    for (uint32_t i = 0; i < lf.n; ++i) present |= (lf.lvl[i] == o);
    if (present || (full && !(o < lf.lvl[lf.n - 1]))) continue;
    uint32_t pos = full ? lf.n - 1 : lf.n++;
    while (pos > 0 && lf.lvl[pos - 1] > o) {
      lf.lvl[pos] = lf.lvl[pos - 1];
      --pos;
    }
    lf.lvl[pos] = o;
  }
  return pass;
}

// Next chunk of a persistent launch's work list for this wavefront: walks the items in
// order and claims one chunk of the first item that has any left.  Returns false when the
// list is drained; `it` carries the wavefront's position between calls.
__device__ __forceinline__ bool claim_chunk(const BatchParams &bp, uint32_t num_items,
                                            uint32_t &it, int lane, uint32_t &job,
                                            uint64_t &cbeg, uint64_t &cend) {
  while (it < num_items) {
    uint32_t c = 0;
    if (lane == 0) c = atomicAdd(bp.claims + it, 1u);
    c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
    const WorkItem wi = bp.items[it];
    const uint64_t b = wi.begin + static_cast<uint64_t>(c) * wi.chunk;
    if (b < wi.end) {
      job = __builtin_amdgcn_readfirstlane(wi.job);
      cbeg = b;
      cend = b + wi.chunk < wi.end ? b + wi.chunk : wi.end;
      return true;
    }
    ++it;
  }
  return false;
}

}  // namespace
}  // namespace sqsg
