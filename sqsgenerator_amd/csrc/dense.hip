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
//
// The MFMA part of a shell is straight-line code over the TT x TT upper tile grid (TT is
// the kernel's tile bound, T <= TT the problem's tile count): every tile load is issued a
// fixed number of grid positions ahead into a register ring and every MFMA runs; tiles
// past the problem's T x T grid read zeros (bounds-checked buffer loads) and zero tiles of
// the grid are stored as zeros, so they add nothing.  Without branches the scheduler can
// overlap the row-end packing (accumulator reads, byte packing) with the next row's MFMAs.
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

// read-only problem data through the constant address space: uniform loads become scalar
// loads, which do not wait behind the tile stream's vector loads
template <class T> __device__ __forceinline__ T ldc(const T *p, uint32_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T &constant_ref(const T *p,
                                                                                   uint32_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}


// launch shapes: tile bound TT, 32-column tile groups G per wavefront (X holds G * TT
// one-hot tiles in registers), resident wavefronts W per SIMD the registers are budgeted
// for, and the tile-stream prefetch depth D (grid positions; it must divide the
// TT (TT + 1) / 2 positions of a shell so every shell reuses the same ring slots).  With
// two wavefronts per SIMD (256 registers each), one generates candidates or sums
// objectives while the other keeps the matrix core busy; large tile bounds need the whole
// register file for the one-hot tiles and the ring.
template <int TT_, int G_, int W_, int D_> struct DenseShape {
  static constexpr int TT = TT_, G = G_, W = W_, D = D_;
  static_assert((TT * (TT + 1) / 2) % D == 0, "ring depth must divide a shell's positions");
};
template <class F> hipError_t with_dense_shape(uint32_t tile_bound, F &&f) {
  switch (tile_bound) {
    case 1: return f(DenseShape<1, 4, 2, 1>{});
    case 2: return f(DenseShape<2, 4, 2, 3>{});
    case 4: return f(DenseShape<4, 2, 2, 5>{});
    case 8: return f(DenseShape<8, 2, 2, 12>{});
    case 16: return f(DenseShape<16, 2, 1, 8>{});
    case 28: return f(DenseShape<28, 1, 1, 14>{});
    case 32: return f(DenseShape<32, 1, 1, 16>{});
    default: return hipErrorInvalidValue;
  }
}

// byte offset of upper tile (I, J) of shell s in the stream (every I <= J < T tile is
// stored, zero tiles included), or an offset past the end (the load returns zeros)
__device__ __forceinline__ uint32_t tile_offset(uint32_t s, uint32_t I, uint32_t J, uint32_t S,
                                                uint32_t T) {
  if (s >= S || I >= T || J >= T) return 0x7fff0000u;
  const uint32_t per_shell = T * (T + 1) / 2;
  return (s * per_shell + I * T - I * (I - 1) / 2 + (J - I)) * 1024u;
}

// Scores the candidates in LDS rows [0, npass * CW) (row c at rows + c * stride, 32 * dT
// bytes, padding sites 0xff); `cnt` is CW * K * K words of LDS.  Passes whose bit in `live`
// is clear are skipped.  Returns the objective of the lane's own candidate
// (lane < npass * CW) and, when given, writes its ordered bonds and SRO cube.
template <int TT, int G, int D, class PR>
__device__ __forceinline__ double dense_score(const PR &pr, const uint8_t *rows,
                                              uint32_t stride, uint32_t *cnt, double *vals,
                                              uint32_t CW, uint32_t npass,
                                              unsigned long long live, int lane, int32_t *bonds,
                                              double *sro, uint32_t out_rows) {
  constexpr int P = TT * (TT + 1) / 2;
  const uint32_t K = pr.K, KK = K * K, T = pr.dT, S = pr.S;
  const uint32_t g = CW / G;
  const uint32_t h = static_cast<uint32_t>(lane) >> 5, n = static_cast<uint32_t>(lane) & 31u;
  const uint32_t cl = n / K, b = n - cl * K;
  const bool col_ok = cl < g;
  const uint32_t bb = b * 0x01010101u;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t *>(pr.dtiles), static_cast<short>(0),
      static_cast<int>(pr.dtiles_n * 1024u), 0x00020000);
  const uint32_t toff = static_cast<uint32_t>(lane) * 16u;
  const uint32_t my_pass = static_cast<uint32_t>(lane) / CW;
  const uint32_t my_cc = static_cast<uint32_t>(lane) - my_pass * CW;  // candidate in its pass
  // objective terms are evaluated lane-parallel: lane l takes terms idx = l, l + 64, ... of
  // the pass (idx = candidate * Ts + term), stepped without division
  const uint32_t Ts = K * (K + 1) / 2, nterm = CW * Ts;
  const uint32_t step_t = 64u % Ts, step_c = 64u / Ts;
  const uint32_t c_first = static_cast<uint32_t>(lane) / Ts;
  const uint32_t t_first = static_cast<uint32_t>(lane) - c_first * Ts;
  const uint32_t cube = pr.S * KK;
  uint32_t keep = 0;  // accumulator registers of GEMM2 on this lane's diagonal block
#pragma unroll
  for (uint32_t r = 0; r < 16; ++r) {
    const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
    if (col_ok && m / K == cl) keep |= 1u << r;
  }
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
    // ring slot q % D holds grid position q (row-major over I <= J < TT) of the shell.
    // Grid positions are stepped incrementally so that every index folds to a constant
    // once the loops are unrolled.
    v4i ring[D];
    {
      int I2 = 0, J2 = 0;
#pragma unroll
      for (int q = 0; q < D; ++q) {
        ring[q] = __builtin_bit_cast(
            v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, tile_offset(0, I2, J2, S, T) + toff,
                                                       0, 0));
        if (++J2 == TT) ++I2, J2 = I2;
      }
    }
    for (uint32_t s = 0; s < S; ++s) {
      v16i C[G], Y[G];
#pragma unroll
      for (int gi = 0; gi < G; ++gi) C[gi] = v16i{};
      int I = 0, J = 0;             // current grid position
      int I2 = 0, J2 = 0, s2 = 0;   // the position D ahead (s2 = 1: next shell)
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (++J2 == TT) {
          ++I2, J2 = I2;
          if (I2 == TT) I2 = 0, J2 = 0, s2 = 1;
        }
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (J == I) {
#pragma unroll
          for (int gi = 0; gi < G; ++gi) Y[gi] = v16i{};
        }
        const v4i u = ring[q % D];
        ring[q % D] = __builtin_bit_cast(
            v4i, __builtin_amdgcn_raw_buffer_load_b128(
                     rs, tile_offset(s + s2, I2, J2, S, T) + toff, 0, 0));
        if (++J2 == TT) {
          ++I2, J2 = I2;
          if (I2 == TT) I2 = 0, J2 = 0, s2 = 1;
        }
#pragma unroll
        for (int gi = 0; gi < G; ++gi)
          Y[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(u, X[gi][J], Y[gi], 0, 0, 0);
        if (J == TT - 1) {  // row I complete: C += X_I^T Y_I
#pragma unroll
          for (int gi = 0; gi < G; ++gi)
            C[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[gi][I], pack_counts(Y[gi]), C[gi],
                                                          0, 0, 0);
        }
        if (++J == TT) ++I, J = I;
      }
      // diagonal blocks -> counts [candidate of the pass][a][b]: register r of lane l is
      // row m = 8(r/4) + 4h + r%4 = cK + a of column n = cK + b
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r)
          if ((keep >> r) & 1u) {
            const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
            cnt[gi * g * KK + K * m + b] = static_cast<uint32_t>(C[gi][r]);
          }
      wave_sync();
      // this shell's terms (optimization.h:85-100) for every candidate of the pass
      {
        const TermParam *tp = pr.tparams + s * Ts;
        uint32_t c = c_first, t = t_first;
        for (uint32_t idx = static_cast<uint32_t>(lane); idx < nterm; idx += 64) {
          const TermParam e = tp[t];
          const uint32_t *cc = cnt + c * KK;
          const uint32_t cab = cc[e.xi * K + e.eta];
          const uint32_t cba = e.xi == e.eta ? 0u : cc[e.eta * K + e.xi];
          const double v = 1.0 - static_cast<double>(cab + cba) * e.pref;
          vals[idx] = e.pw * absolute(v - e.tgt);
          const uint32_t cr = pass * CW + c;  // candidate row
          if ((bonds || sro) && cr < out_rows) {
            const uint32_t o = cr * cube + (s * K + e.xi) * K + e.eta;
            const uint32_t ot = cr * cube + (s * K + e.eta) * K + e.xi;
            if (sro) sro[o] = v, sro[ot] = v;
            if (bonds) bonds[o] = static_cast<int32_t>(cab), bonds[ot] = static_cast<int32_t>(cc[e.eta * K + e.xi]);
          }
          t += step_t;
          c += step_c;
          if (t >= Ts) t -= Ts, ++c;
        }
      }
      wave_sync();
      // each candidate's lane adds the shell's terms in the reference order
      if (my_pass == pass) {
        const double *vv = vals + my_cc * Ts;
        for (uint32_t t = 0; t < Ts; ++t) obj += vv[t];
      }
      wave_sync();
    }
  }
  return obj;
}

template <class PR> __device__ __forceinline__ void load_dense_row(uint8_t *row, const PR &pr) {
  const uint32_t *src = reinterpret_cast<const uint32_t *>(pr.species0D);
  uint32_t *dst = reinterpret_cast<uint32_t *>(row);
  for (uint32_t w = 0; w < 8 * pr.dT; ++w) dst[w] = src[w];
}

// One segment of a job: gen_group lanes x p.rounds candidates starting at sbeg (the
// candidate numbering of the lane kernel with 64 replaced by gen_group); candidates at or
// beyond `end` are inactive.
template <int TT, int G, int D, class RP>
__device__ __forceinline__ void dense_segment(const RP &p, uint64_t sbeg, uint64_t end,
                                              uint8_t *wl, int lane, LevelFilter &lf,
                                              unsigned long long &scored) {
  const auto &pr = p.prob[0];
  const uint32_t GEN = p.gen_group, CW = p.dense_cw, stride = p.dense_stride;
  const uint32_t npass = GEN / CW;
  uint8_t *rows = wl + p.off_rows[0];
  uint32_t *cnt = reinterpret_cast<uint32_t *>(wl + p.off_hist);
  double *vals = reinterpret_cast<double *>(wl + p.off_obj);
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
        dense_score<TT, G, D>(pr, rows, stride, cnt, vals, CW, npass, live, lane, nullptr, nullptr,
                              0u);
    scored += static_cast<unsigned long long>(__popcll(am));
    const unsigned long long g =
        __hip_atomic_load(p.thr_key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the wave's own L-th distinct level bounds the final levels as well: a candidate above
    // it has L better distinct objectives and cannot be kept, so it is not recorded
    const bool below = active && okey(tot) <= g;
    offer_levels(lf, p, __ballot(below), tot, g, lane);
    const unsigned long long lk = lf.n == p.L ? okey(lf.lvl[p.L - 1]) : ~0ull;
    const bool pass = below && okey(tot) <= lk;
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
    wave_sync();  // rows are rewritten by the next round
  }
}

template <int TT, int G, int W, int D>
__global__ __launch_bounds__(64, W) void run_dense_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const uint64_t per_seg = static_cast<uint64_t>(p.gen_group) * p.rounds;
  const uint64_t total = p.cand_end - p.cand_begin;
  const uint64_t nseg = (total + per_seg - 1) / per_seg;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  // segments are pulled from a counter: wavefronts that start late (more blocks than fit
  // at once) or run slow do not hold up the launch
  for (;;) {
    uint32_t seg = 0;
    if (lane == 0) seg = atomicAdd(p.next_seg, 1u);
    seg = __builtin_amdgcn_readfirstlane(__shfl(seg, 0));
    if (seg >= nseg) break;
    dense_segment<TT, G, D>(p, p.cand_begin + seg * per_seg, p.cand_end, lds, lane, lf, scored);
  }
  if (lane == 0 && scored) atomicAdd(p.scored, scored);
}

// Persistent form: every wavefront claims one-segment chunks of the jobs' candidate ranges
// until the work list is drained; the jobs of a launch share the tile bound TT.
template <int TT, int G, int W, int D>
__global__ __launch_bounds__(64, W) void batch_dense_kernel(BatchParams bp) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const uint32_t num_items = *bp.num_items;
  uint32_t last_job = 0xffffffffu, it = 0, job = 0;
  uint64_t cbeg = 0, cend = 0;
  LevelFilter lf;
  lf.n = 0;
  unsigned long long scored = 0;
  while (claim_chunk(bp, num_items, it, lane, job, cbeg, cend)) {
    // the job table is read-only during the launch: scalar loads through the constant
    // address space keep every job parameter uniform
    const auto &p = constant_ref(bp.jobs, job);
    if (job != last_job) {
      if (last_job != 0xffffffffu && lane == 0 && scored)
        atomicAdd(bp.jobs[last_job].scored, scored);
      scored = 0;
      lf.n = 0;
      last_job = job;
    }
    dense_segment<TT, G, D>(p, cbeg, cend, lds, lane, lf, scored);
  }
  if (last_job != 0xffffffffu && lane == 0 && scored) atomicAdd(bp.jobs[last_job].scored, scored);
}

template <int TT, int G, int D>
__global__ __launch_bounds__(64) void score_dense_kernel(ScoreParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  const DevProblem &pr = p.prob;
  const uint32_t GEN = p.dense_gen, CW = p.dense_cw, npass = GEN / CW;
  uint8_t *rows = lds + p.off_row;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(lds + p.off_hist);
  double *vals = reinterpret_cast<double *>(lds + p.off_terms);
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
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * GEN;
  const uint32_t out_rows = static_cast<uint32_t>(p.B - b0 < GEN ? p.B - b0 : GEN);
  const double obj = dense_score<TT, G, D>(pr, rows, p.dense_stride, cnt, vals, CW, npass, live,
                                           lane, p.bonds ? p.bonds + b0 * nb : nullptr,
                                           p.sro ? p.sro + b0 * nb : nullptr, out_rows);
  if (active) p.objective[b] = obj;
}

hipError_t dense_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

template <class Sh> hipError_t run_dense_k(const RunParams &p, uint32_t blocks, hipStream_t s) {
  const void *fn = reinterpret_cast<const void *>(&run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>);
  const hipError_t e = dense_lds(fn, p.wave_lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>), dim3(blocks), dim3(64),
                     p.wave_lds, s, p);
  return hipGetLastError();
}

template <class Sh> hipError_t batch_dense_k(const BatchParams &bp, uint32_t blocks,
                                              hipStream_t s) {
  hipLaunchKernelGGL((batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>), dim3(blocks), dim3(64),
                     bp.wave_lds, s, bp);
  return hipGetLastError();
}

template <class Sh> hipError_t score_dense_k(const ScoreParams &p, hipStream_t s) {
  const void *fn = reinterpret_cast<const void *>(&score_dense_kernel<Sh::TT, Sh::G, Sh::D>);
  const hipError_t e = dense_lds(fn, p.wave_lds);
  if (e != hipSuccess) return e;
  const uint32_t blocks = static_cast<uint32_t>((p.B + p.dense_gen - 1) / p.dense_gen);
  hipLaunchKernelGGL((score_dense_kernel<Sh::TT, Sh::G, Sh::D>), dim3(blocks), dim3(64),
                     p.wave_lds, s, p);
  return hipGetLastError();
}

}  // namespace

uint32_t dense_tile_bound(uint32_t tiles) {
  for (uint32_t tt : {1u, 2u, 4u, 8u, 16u, 28u, 32u})
    if (tiles <= tt) return tt;
  return 0;
}
uint32_t dense_groups(uint32_t tile_bound) {
  uint32_t g = 0;
  (void)with_dense_shape(tile_bound, [&](auto sh) {
    g = decltype(sh)::G;
    return hipSuccess;
  });
  return g;
}

uint32_t dense_waves_per_cu(uint32_t tile_bound, size_t wave_lds, bool persistent) {
  int n = 0;
  const hipError_t e = with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    const void *fn = persistent
                         ? reinterpret_cast<const void *>(&batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>)
                         : reinterpret_cast<const void *>(&run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>);
    hipError_t r = dense_lds(fn, wave_lds);
    if (r != hipSuccess) return r;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, fn, 64, wave_lds);
  });
  return e == hipSuccess && n > 0 ? static_cast<uint32_t>(n) : 1u;
}

hipError_t launch_run_dense(const RunParams &p, uint32_t tile_bound, uint32_t blocks,
                            hipStream_t s) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return run_dense_k<Sh>(p, blocks, s);
  });
}

hipError_t prepare_batch_dense(uint32_t tile_bound, size_t lds) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return dense_lds(
        reinterpret_cast<const void *>(&batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D>), lds);
  });
}

hipError_t launch_batch_dense(const BatchParams &bp, uint32_t tile_bound, uint32_t blocks,
                              hipStream_t s) {
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return batch_dense_k<Sh>(bp, blocks, s);
  });
}

hipError_t launch_score_dense(const ScoreParams &p, uint32_t tile_bound, hipStream_t s) {
  if (p.B == 0) return hipSuccess;
  return with_dense_shape(tile_bound, [&](auto sh) {
    using Sh = decltype(sh);
    return score_dense_k<Sh>(p, s);
  });
}

}  // namespace sqsg
