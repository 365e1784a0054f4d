// Exact dense bond counting on the MFMA cores of gfx950 (v_mfma_i32_32x32x32_i8).
//
// For shell s let U_s be the strictly upper shell adjacency of the sorted sites
// (U_s[i][j] = 1 when i < j are s-th neighbours) and X the one-hot species matrix of a
// candidate (X[j][b] = 1 when site j holds species b).  Then
//     bonds(s, a, b) = sum_i X[i][a] * (U_s X)[i][b]              (optimization.h:74-77)
// exactly, in int32.  A wavefront scores CW = G * g candidates per pass, g = 32 / K of them
// side by side in the 32 columns of a tile group (column n = candidate n / K, species n % K):
//   GEMM1  Y_I = sum_{J >= I} U_s[I][J] * X_J       A = adjacency tile, B = one-hot tile
//   GEMM2  C  += X_I^T * Y_I                        A = one-hot tile,  B = Y_I as int8
// C's diagonal K x K blocks are the candidates' bond counts.  The operand layouts (A: lane
// holds row lane % 32, B: lane holds column lane % 32, bytes q of half h = lane / 32 share
// reduction index 16h + q; D: register r of lane l is row 8(r/4) + 4h + r%4, column l % 32)
// make both GEMMs meet without data movement:
//  * one-hot tile J of a lane = bytes [species(32J + 16h + q) == b], q = 0..15; the same
//    register is GEMM1's B operand for column (c, b) and GEMM2's A operand for row (c, a);
//  * adjacency tile rows are stored permuted (row m holds site 16h + 4(m/8) + m%4 with
//    h = (m/4)%2), so accumulator register r of lane l holds Y at site 32I + 16h + r and
//    the 16 registers pack straight into GEMM2's B operand (Y <= 127 is checked on the host).
// Shells are processed in order and each shell's objective terms are summed as soon as its
// counts are out, which keeps the FP64 summation order of optimization.h:85-100.
#include <hip/hip_runtime.h>

#include "candidates.hpp"
#include "kernels.hpp"

namespace sqsg {
namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// bytes of w equal to the species replicated in bb -> 1, others -> 0.  Bytes are below
// 0x80 or the 0xff padding; (x | 0x80) - 1 keeps bit 7 unless x == 0, and no byte borrows.
__device__ __forceinline__ int onehot4(uint32_t w, uint32_t bb) {
  const uint32_t x = w ^ bb;
  const uint32_t t = (x | 0x80808080u) - 0x01010101u;
  return static_cast<int>((~t >> 7) & 0x01010101u);
}

// GEMM1 accumulator (16 counts <= 127 of one column) -> GEMM2 B operand bytes
__device__ __forceinline__ v4i pack_counts(const v16i &y) {
  v4i b;
#pragma unroll
  for (int w = 0; w < 4; ++w)
    b[w] = y[4 * w] | (y[4 * w + 1] << 8) | (y[4 * w + 2] << 16) | (y[4 * w + 3] << 24);
  return b;
}

__device__ __forceinline__ unsigned long long low_mask(uint32_t n) {
  return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

// Scores the candidates in LDS rows [0, npass * CW) (row c at rows + c * stride, 32 * dT
// bytes, padding sites 0xff).  Passes whose bit in `live` is clear are skipped.  Returns
// the objective of the lane's own candidate (lane < npass * CW) and, when given, writes
// its ordered bonds and SRO cube.
template <int TT, int G>
__device__ __forceinline__ double dense_score(const DevProblem &pr, const uint8_t *rows, uint32_t stride,
                              uint32_t *cnt, uint32_t CW, uint32_t npass,
                              unsigned long long live, int lane, int32_t *bonds, double *sro) {
  const uint32_t K = pr.K, KK = K * K, T = pr.dT;
  const uint32_t g = CW / G;
  const uint32_t h = static_cast<uint32_t>(lane) >> 5, n = static_cast<uint32_t>(lane) & 31u;
  const uint32_t cl = n / K, b = n - cl * K;
  const bool col_ok = cl < g;
  const uint32_t bb = b * 0x01010101u;
  uint32_t keep = 0;  // accumulator registers of GEMM2 on this lane's diagonal block
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {
    const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
    if (col_ok && m / K == cl) keep |= 1u << r;
  }
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t *>(pr.dtiles), static_cast<short>(0),
      static_cast<int>(pr.dtiles_n * 1024u), 0x00020000);
  const uint32_t toff = static_cast<uint32_t>(lane) * 16u;
  const uint32_t my_pass = static_cast<uint32_t>(lane) / CW;
  double obj = 0.0;
  for (uint32_t pass = 0; pass < npass; ++pass) {
    if (!((live >> pass) & 1ull)) continue;
    v4i X[G][TT];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const uint32_t c = pass * CW + gi * g + cl;
      const uint8_t *row = rows + (col_ok ? c : 0u) * stride + 16 * h;
#pragma unroll
      for (int J = 0; J < TT; ++J) {
        v4i x = {0, 0, 0, 0};
        if (static_cast<uint32_t>(J) < T && col_ok) {
          const uint4 w = *reinterpret_cast<const uint4 *>(row + 32 * J);
          x[0] = onehot4(w.x, bb);
          x[1] = onehot4(w.y, bb);
          x[2] = onehot4(w.z, bb);
          x[3] = onehot4(w.w, bb);
        }
        X[gi][J] = x;
      }
    }
    uint32_t t = 0;  // position in the tile stream
    v4i cur = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, toff, 0, 0));
    v4i nxt = __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, toff + 1024u, 0, 0));
    for (uint32_t s = 0; s < pr.S; ++s) {
      v16i C[G];
#pragma unroll
      for (int gi = 0; gi < G; ++gi) C[gi] = v16i{};
#pragma unroll
      for (int I = 0; I < TT; ++I) {
        if (static_cast<uint32_t>(I) >= T) continue;
        const uint32_t mask = pr.dmask[s * T + I];  // bits J < T only
        v16i Y[G];
#pragma unroll
        for (int gi = 0; gi < G; ++gi) Y[gi] = v16i{};
#pragma unroll
        for (int J = I; J < TT; ++J) {
          if ((mask >> J) & 1u) {
            const v4i u = cur;
            cur = nxt;
            nxt = __builtin_bit_cast(
                v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, (t + 2) * 1024u + toff, 0, 0));
            ++t;
#pragma unroll
            for (int gi = 0; gi < G; ++gi)
              Y[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(u, X[gi][J], Y[gi], 0, 0, 0);
          }
        }
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
          C[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[gi][I], pack_counts(Y[gi]), C[gi],
                                                        0, 0, 0);
      }
      // diagonal blocks -> counts [candidate of the pass][a][b]
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r)
          if ((keep >> r) & 1u) {
            const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
            cnt[gi * g * KK + K * m + b] = static_cast<uint32_t>(C[gi][r]);
          }
      wave_sync();
      if (my_pass == pass) {
        const uint32_t *cc = cnt + (static_cast<uint32_t>(lane) - pass * CW) * KK;
        const double *pref = pr.pref + s * KK;
        const double *pw = pr.pw + s * KK;
        const double *tgt = pr.tgt + s * KK;
        for (uint32_t xi = 0; xi < K; ++xi) {
          const uint32_t cxx = cc[xi * K + xi];
          const double sxx = 1.0 - static_cast<double>(cxx) * pref[xi * K + xi];
          obj += pw[xi * K + xi] * absolute(sxx - tgt[xi * K + xi]);
          if (sro) sro[(s * K + xi) * K + xi] = sxx;
          if (bonds) bonds[(s * K + xi) * K + xi] = static_cast<int32_t>(cxx);
          for (uint32_t eta = xi + 1; eta < K; ++eta) {
            const uint32_t cab = cc[xi * K + eta], cba = cc[eta * K + xi];
            const double v = 1.0 - static_cast<double>(cab + cba) * pref[xi * K + eta];
            obj += pw[xi * K + eta] * absolute(v - tgt[xi * K + eta]);
            if (sro) sro[(s * K + xi) * K + eta] = v, sro[(s * K + eta) * K + xi] = v;
            if (bonds) {
              bonds[(s * K + xi) * K + eta] = static_cast<int32_t>(cab);
              bonds[(s * K + eta) * K + xi] = static_cast<int32_t>(cba);
            }
          }
        }
      }
      wave_sync();
    }
  }
  return obj;
}

__device__ __forceinline__ void load_dense_row(uint8_t *row, const DevProblem &pr) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(pr.species0D);
  uint32_t *dst = reinterpret_cast<uint32_t *>(row);
  for (uint32_t w = 0; w < 8 * pr.dT; ++w) dst[w] = src[w];
}

// One segment of a job: gen_group lanes x p.rounds candidates starting at sbeg (the
// candidate numbering of the lane kernel with 64 replaced by gen_group); candidates at or
// beyond `end` are inactive.
template <int TT, int G>
__device__ __forceinline__ void dense_segment(const RunParams &p, uint64_t sbeg, uint64_t end, uint8_t *wl,
                              int lane, LevelFilter &lf, unsigned long long &scored) {
  const DevProblem &pr = p.prob[0];
  const uint32_t GEN = p.gen_group, CW = p.dense_cw, stride = p.dense_stride;
  const uint32_t npass = GEN / CW;
  uint8_t *rows = wl + p.off_rows[0];
  uint32_t *cnt = reinterpret_cast<uint32_t *>(wl + p.off_hist);
  const bool gl = static_cast<uint32_t>(lane) < GEN;
  uint8_t *row = rows + (gl ? static_cast<uint32_t>(lane) : 0u) * stride;
  for (uint32_t r = 0; r < p.rounds; ++r) {
    const uint64_t cand = p.systematic ? sbeg + static_cast<uint64_t>(lane) * p.rounds + r
                                       : sbeg + static_cast<uint64_t>(r) * GEN + lane;
    const bool active = gl && cand < end;
    if (gl) {
      if (p.systematic) {
        if (r == 0) {
          load_dense_row(row, pr);
          if (active) unrank_candidate(Strided<0>{row}, pr, cand + 1, 1u);
        } else if (active) {
          next_perm_candidate(Strided<0>{row}, pr.bounds[0], pr.bounds[1] - pr.bounds[0]);
        }
      } else {
        load_dense_row(row, pr);
        if (active && !(p.chunk_quirk && cand % p.chunk_size == 0))
          shuffle_candidate(Strided<0>{row}, pr, cand);
      }
    }
    const unsigned long long am = __ballot(active);
    if (am == 0) break;  // the tail of a systematic segment
    unsigned long long live = 0;
    for (uint32_t q = 0; q < npass; ++q)
      if ((am >> (q * CW)) & low_mask(CW)) live |= 1ull << q;
    wave_sync();
    const double tot =
        dense_score<TT, G>(pr, rows, stride, cnt, CW, npass, live, lane, nullptr, nullptr);
    scored += static_cast<unsigned long long>(__popcll(am));
    const unsigned long long g =
        __hip_atomic_load(p.thr_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const bool pass = active && okey(tot) <= g;
    if (pass) {
      const uint32_t slot = atomicAdd(p.rec_count, 1u);
      if (slot < p.rec_cap) {
        uint8_t *dst = p.rec_species + static_cast<size_t>(slot) * p.rec_stride;
        for (uint32_t i = 0; i < pr.N; ++i) dst[i] = row[i];
        p.rec_cand[slot] = cand;
        p.rec_obj[slot] = tot;
        p.rec_sub[slot] = tot;
      }
    }
    offer_levels(lf, p, __ballot(pass), tot, g, lane);
    wave_sync();  // rows are rewritten by the next round
  }
}

template <int TT, int G>
__global__ __launch_bounds__(64) void run_dense_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const uint64_t per_seg = static_cast<uint64_t>(p.gen_group) * p.rounds;
  const uint64_t total = p.cand_end - p.cand_begin;
  const uint64_t nseg = (total + per_seg - 1) / per_seg;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  for (uint64_t seg = blockIdx.x; seg < nseg; seg += gridDim.x)
    dense_segment<TT, G>(p, p.cand_begin + seg * per_seg, p.cand_end, lds, lane, lf, scored);
  if (lane == 0 && scored) atomicAdd(p.scored, scored);
}

// Persistent form: every wavefront pulls work items (job, candidate range) from a shared
// counter until the list is exhausted; the jobs of a launch share the tile bound TT.
template <int TT, int G>
__global__ __launch_bounds__(64) void batch_dense_kernel(BatchParams bp) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const uint32_t num_items = *bp.num_items;
  uint32_t last_job = 0xffffffffu;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  for (;;) {
    uint32_t item = 0;
    if (lane == 0) item = atomicAdd(bp.next_item, 1u);
    item = __builtin_amdgcn_readfirstlane(__shfl(item, 0));
    if (item >= num_items) break;  // every wave leaves once the list is drained
    const WorkItem wi = bp.items[item];
    const RunParams &p = bp.jobs[wi.job];
    if (wi.job != last_job) {
      if (last_job != 0xffffffffu && lane == 0 && scored)
        atomicAdd(bp.jobs[last_job].scored, scored);
      scored = 0;
      lf.n = 0;
      last_job = wi.job;
    }
    const uint64_t per_seg = static_cast<uint64_t>(p.gen_group) * p.rounds;
    for (uint64_t sbeg = wi.begin; sbeg < wi.end; sbeg += per_seg)
      dense_segment<TT, G>(p, sbeg, wi.end, lds, lane, lf, scored);
  }
  if (last_job != 0xffffffffu && lane == 0 && scored) atomicAdd(bp.jobs[last_job].scored, scored);
}

template <int TT, int G>
__global__ __launch_bounds__(64) void score_dense_kernel(ScoreParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const DevProblem &pr = p.prob;
  const uint32_t GEN = p.dense_gen, CW = p.dense_cw, npass = GEN / CW;
  uint8_t *rows = lds + p.off_row;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(lds + p.off_hist);
  const uint32_t nb = pr.S * pr.K * pr.K;
  const uint64_t b = static_cast<uint64_t>(blockIdx.x) * GEN + lane;
  const bool gl = static_cast<uint32_t>(lane) < GEN;
  const bool active = gl && b < p.B;
  if (gl) {
    uint8_t *row = rows + lane * p.dense_stride;
    for (uint32_t i = 0; i < 32 * pr.dT; ++i)
      row[i] = (active && i < pr.N) ? p.species[b * pr.N + i] : (i < pr.N ? 0 : 0xff);
  }
  const unsigned long long am = __ballot(active);
  unsigned long long live = 0;
  for (uint32_t q = 0; q < npass; ++q)
    if ((am >> (q * CW)) & low_mask(CW)) live |= 1ull << q;
  wave_sync();
  const double obj = dense_score<TT, G>(pr, rows, p.dense_stride, cnt, CW, npass, live, lane,
                                        active && p.bonds ? p.bonds + b * nb : nullptr,
                                        active && p.sro ? p.sro + b * nb : nullptr);
  if (active) p.objective[b] = obj;
}

hipError_t dense_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

template <int TT, int G> hipError_t run_dense_k(const RunParams &p, uint32_t blocks,
                                                hipStream_t s) {
  const hipError_t e = dense_lds(reinterpret_cast<const void *>(&run_dense_kernel<TT, G>),
                                 p.wave_lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((run_dense_kernel<TT, G>), dim3(blocks), dim3(64), p.wave_lds, s, p);
  return hipGetLastError();
}

template <int TT, int G> hipError_t batch_dense_k(const BatchParams &bp, uint32_t blocks,
                                                  hipStream_t s) {
  hipLaunchKernelGGL((batch_dense_kernel<TT, G>), dim3(blocks), dim3(64), bp.wave_lds, s, bp);
  return hipGetLastError();
}

template <int TT, int G> hipError_t score_dense_k(const ScoreParams &p, hipStream_t s) {
  const hipError_t e = dense_lds(reinterpret_cast<const void *>(&score_dense_kernel<TT, G>),
                                 p.wave_lds);
  if (e != hipSuccess) return e;
  const uint32_t blocks = static_cast<uint32_t>((p.B + p.dense_gen - 1) / p.dense_gen);
  hipLaunchKernelGGL((score_dense_kernel<TT, G>), dim3(blocks), dim3(64), p.wave_lds, s, p);
  return hipGetLastError();
}

}  // namespace

uint32_t dense_tile_bound(uint32_t tiles) { return tiles <= 8 ? 8 : tiles <= 16 ? 16 : 32; }
uint32_t dense_groups(uint32_t tile_bound) { return tile_bound == 8 ? 4 : tile_bound == 16 ? 2 : 1; }

hipError_t launch_run_dense(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                            hipStream_t s) {
  switch (tile_bound) {
    case 8: return run_dense_k<8, 4>(p, blocks, s);
    case 16: return run_dense_k<16, 2>(p, blocks, s);
    case 32: return run_dense_k<32, 1>(p, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t prepare_batch_dense(uint32_t tile_bound, size_t lds) {
  switch (tile_bound) {
    case 8: return dense_lds(reinterpret_cast<const void *>(&batch_dense_kernel<8, 4>), lds);
    case 16: return dense_lds(reinterpret_cast<const void *>(&batch_dense_kernel<16, 2>), lds);
    case 32: return dense_lds(reinterpret_cast<const void *>(&batch_dense_kernel<32, 1>), lds);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_batch_dense(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                              hipStream_t s) {
  switch (tile_bound) {
    case 8: return batch_dense_k<8, 4>(bp, blocks, s);
    case 16: return batch_dense_k<16, 2>(bp, blocks, s);
    case 32: return batch_dense_k<32, 1>(bp, blocks, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_score_dense(const ScoreParams &p, uint32_t tile_bound, hipStream_t s) {
  if (p.B == 0) return hipSuccess;
  switch (tile_bound) {
    case 8: return score_dense_k<8, 4>(p, s);
    case 16: return score_dense_k<16, 2>(p, s);
    case 32: return score_dense_k<32, 1>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace sqsg
