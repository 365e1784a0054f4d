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
// fixed number of grid positions ahead into a register ring and every MFMA runs; the tile
// stream holds the whole TT x TT upper grid, tiles past the problem's T x T grid and zero
// tiles of it stored as zeros, so they add nothing.  Without branches the scheduler can
// overlap the row-end packing (accumulator reads, byte packing) with the next row's MFMAs.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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

// One-hot tiles of a pass: X[gi][J] of lane (h, n) holds, for its column's candidate
// first + gi * g + cl (row at rows + c * stride) and species b (bb = b replicated), bytes
// [species(32J + 16h + q) == b], q = 0..15; zero for tiles J >= T and for columns past the
// group's candidates (col_ok false).  BATCHED issues every row read before the first
// conversion (tiles past T read tile 0 and are zeroed): one LDS round trip instead of one per
// tile, where the registers allow it.
template <int TT, int G, bool BATCHED>
__device__ __forceinline__ void onehot_tiles(v4i (&X)[G][TT], const uint8_t *rows,
                                             uint32_t first, uint32_t g, uint32_t cl,
                                             bool col_ok, uint32_t stride, uint32_t h,
                                             uint32_t T, uint32_t bb) {
  if constexpr (BATCHED) {
    uint4 raw[G][TT];
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const uint8_t *row = rows + (col_ok ? first + gi * g + cl : 0u) * stride + 16 * h;
#pragma unroll
      for (int J = 0; J < TT; ++J)
        raw[gi][J] =
            *reinterpret_cast<const uint4 *>(row + 32 * (static_cast<uint32_t>(J) < T ? J : 0));
    }
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
#pragma unroll
      for (int J = 0; J < TT; ++J) {
        const bool on = static_cast<uint32_t>(J) < T && col_ok;
        const uint4 w = raw[gi][J];
        v4i x;
        x[0] = on ? onehot4(w.x, bb) : 0;
        x[1] = on ? onehot4(w.y, bb) : 0;
        x[2] = on ? onehot4(w.z, bb) : 0;
        x[3] = on ? onehot4(w.w, bb) : 0;
        X[gi][J] = x;
      }
  } else {
#pragma unroll
    for (int gi = 0; gi < G; ++gi) {
      const uint8_t *row = rows + (col_ok ? first + gi * g + cl : 0u) * stride + 16 * h;
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
  }
}

// GEMM1 accumulator (16 counts <= 127 of one column) -> GEMM2 B operand bytes.  Per word
// three byte permutes (v_perm_b32: selector bytes 0-3 take bytes of the second operand, 4-7
// of the first, 0x0c gives zero): [y0, y1, 0, 0] and [0, 0, y2, y3], then their bytes
// merged; the shift-or form of the same packing compiles to four operations per word.
__device__ __forceinline__ v4i pack_counts(const v16i &y) {
  v4i b;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t lo = __builtin_amdgcn_perm(static_cast<uint32_t>(y[4 * w + 1]),
                                              static_cast<uint32_t>(y[4 * w]), 0x0c0c0400u);
    const uint32_t hi = __builtin_amdgcn_perm(static_cast<uint32_t>(y[4 * w + 3]),
                                              static_cast<uint32_t>(y[4 * w + 2]), 0x04000c0cu);
    b[w] = static_cast<int>(__builtin_amdgcn_perm(hi, lo, 0x07060100u));
  }
  return b;
}

__device__ __forceinline__ unsigned long long low_mask(uint32_t n) {
  return n >= 64 ? ~0ull : ((1ull << n) - 1ull);
}

// Mask of the GEMM2 accumulator registers of a lane that fall on its candidate's diagonal
// block.  Register r holds row m = 8(r/4) + 4h + r%4 of the lane's column, and the block is
// rows [cK, cK + K): with `rows` the bit set of those rows, nibble j of the mask is bits
// 8j + 4h .. 8j + 4h + 3 of it.  Shifts only, no division per register.
__device__ __forceinline__ uint32_t diag_registers(uint32_t cK, uint32_t K, uint32_t h,
                                                   bool col_ok) {
  const uint64_t rows = ((1ull << K) - 1ull) << cK;
  uint32_t keep = 0;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j)
    keep |= static_cast<uint32_t>((rows >> (8 * j + 4 * h)) & 0xfull) << (4 * j);
  return col_ok ? keep : 0u;
}

// read-only problem data through the constant address space: uniform loads become scalar
// loads, which do not wait behind the tile stream's vector loads
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T &constant_ref(const T *p,
                                                                                   uint32_t i) {
  return ((const __attribute__((address_space(4))) T *)p)[i];
}


// launch shapes: tile bound TT, 32-column tile groups G per wavefront (X holds G * TT
// one-hot tiles in registers), resident wavefronts W per SIMD the registers are budgeted
// for, and the tile-stream prefetch depth D (grid positions; it must divide the
// TT (TT + 1) / 2 positions of a shell so every shell reuses the same ring slots).  Every
// shape runs two wavefronts per SIMD (256 registers each): while one generates candidates,
// sums objectives or waits on the tile stream, the other keeps the matrix core busy.  For
// the large tile bounds the one-hot tiles take most of the registers, so their rings are
// short (measured: a second wavefront per SIMD beats a deeper ring by 33-58% at 500-1024
// sites).
// Every grid position of the bound costs its MFMAs whatever the problem's tile count, so the
// bounds are spaced closely where cells are common (measured against the next larger bound:
// +26% at 192 sites with 6, +53% at 384 with 12, +27% at 432 with 14, +78% at 600 with 20,
// +27% at 768 with 24, +7% at 864 with 27; a bound of 3 was 35% slower than 4 at 96 sites,
// its four tile groups spill).
// DS is the ring depth of the explicit-scoring and paired kernels, which run with one
// resident wavefront per SIMD where a deeper ring hides the tile loads; the persistent and
// single-job kernels share their SIMD between two wavefronts and keep the registers instead.
template <int TT_, int G_, int W_, int D_, int DS_ = D_> struct DenseShape {
  static constexpr int TT = TT_, G = G_, W = W_, D = D_, DS = DS_;
  static_assert((TT * (TT + 1) / 2) % D == 0 && (TT * (TT + 1) / 2) % DS == 0,
                "ring depth must divide a shell's positions");
};
template <class F> hipError_t with_dense_shape(uint32_t tile_bound, F &&f) {
  switch (tile_bound) {
    case 1: return f(DenseShape<1, 4, 2, 1>{});
    case 2: return f(DenseShape<2, 4, 2, 3>{});
    case 4: return f(DenseShape<4, 2, 2, 5>{});
    case 6: return f(DenseShape<6, 2, 2, 3, 7>{});
    case 8: return f(DenseShape<8, 2, 2, 3, 9>{});
    case 12: return f(DenseShape<12, 2, 2, 3>{});
    case 14: return f(DenseShape<14, 2, 2, 5>{});
    case 16: return f(DenseShape<16, 2, 2, 2>{});
    case 20: return f(DenseShape<20, 1, 2, 7>{});
    case 24: return f(DenseShape<24, 1, 2, 6>{});
    case 27: return f(DenseShape<27, 1, 2, 9>{});
    case 32: return f(DenseShape<32, 1, 2, 4>{});
    default: return hipErrorInvalidValue;
  }
}

// Byte offset in the tile stream of grid position q (row-major over the upper tiles
// I <= J < TT) of shell s.  The stream holds the kernel's whole upper grid for every shell
// (tiles past the problem's T tiles are zeros) and one zero shell after the last, so the
// offset is a constant plus a shell base and every offset the kernels form, the prefetch
// D positions past the last shell's end included, lies inside the stream.  Measured faster
// than a stream of the problem's own T tiles with per-position range conditions at every
// tile bound (+4% at 16 tiles, +7% at 32).
template <int TT>
__device__ __forceinline__ uint32_t tile_offset(uint32_t s, int q) {
  return (s * static_cast<uint32_t>(TT * (TT + 1) / 2) + static_cast<uint32_t>(q)) * 1024u;
}

// Loads one lane's 16 bytes of the tile at stream offset `off` (uniform): the offset goes
// in the instruction's scalar operand and the lane's part in the vector operand, so a load
// costs no vector instruction.  The scalar operand is outside the buffer's range check;
// tile_offset keeps it inside the stream.
__device__ __forceinline__ v4i tile_load(__amdgpu_buffer_rsrc_t rs, uint32_t toff, uint32_t off) {
  return __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rs, toff, off, 0));
}

// One shell's objective terms (optimization.h:85-100) of the candidate whose K x K counts
// are at cc (LDS), added to obj in the reference order (xi, then eta >= xi); bonds and SRO
// cubes [S][K][K] are written when given.  The term parameters are uniform across the
// wavefront and come through scalar loads; with K known at compile time every count read
// has a constant offset.
template <int KT, class PR>
__device__ __forceinline__ void shell_terms_k(const PR &pr, uint32_t s, const uint32_t *cc,
                                              double &obj, int32_t *bonds, double *sro,
                                              uint32_t Kr) {
  const uint32_t K = KT > 0 ? static_cast<uint32_t>(KT) : Kr;
  const TermParam *tp = pr.tparams + s * (K * (K + 1) / 2);
  uint32_t t = 0;
#pragma unroll
  for (uint32_t xi = 0; xi < K; ++xi)
#pragma unroll
    for (uint32_t eta = xi; eta < K; ++eta, ++t) {
      const auto &e = constant_ref(tp, t);
      const uint32_t cab = cc[xi * K + eta];
      const uint32_t cba = xi == eta ? 0u : cc[eta * K + xi];
      const double v = 1.0 - static_cast<double>(cab + cba) * e.pref;
      obj += e.pw * term_abs(v - e.tgt);
      if (sro) sro[(s * K + xi) * K + eta] = v, sro[(s * K + eta) * K + xi] = v;
      if (bonds) {
        bonds[(s * K + xi) * K + eta] = static_cast<int32_t>(cab);
        bonds[(s * K + eta) * K + xi] = static_cast<int32_t>(cc[eta * K + xi]);
      }
    }
}

template <class PR>
__device__ __forceinline__ void shell_terms(const PR &pr, uint32_t s, const uint32_t *cc,
                                            double &obj, int32_t *bonds, double *sro) {
  switch (pr.K) {
    case 2: return shell_terms_k<2>(pr, s, cc, obj, bonds, sro, 2);
    case 3: return shell_terms_k<3>(pr, s, cc, obj, bonds, sro, 3);
    case 4: return shell_terms_k<4>(pr, s, cc, obj, bonds, sro, 4);
    case 5: return shell_terms_k<5>(pr, s, cc, obj, bonds, sro, 5);
    case 6: return shell_terms_k<6>(pr, s, cc, obj, bonds, sro, 6);
    default: return shell_terms_k<0>(pr, s, cc, obj, bonds, sro, pr.K);
  }
}

// Search kernels with one one-hot column fewer than species (DevProblem::Kc = K - 1): the
// last species L = K - 1 has no column, and cc holds the (K-1) x (K-1) counts C[a][b] of the
// others.  With F[a][b] = C[a][b] + C[b][a] (F[a][a] = 2 C[a][a]) the bonds at the sites of
// species a add up to D[a] = sum over those sites of their neighbour counts in the shell,
// which the host has checked to be the same for every candidate (dsum[s][a]).  So
//     c(a, L) = F[a][L] = D[a] - sum_{b < L} F[a][b]
//     c(L, L) = (D[L] - sum_{a < L} F[a][L]) / 2
// exactly, and the terms follow from the same counts as with all K columns, in the same
// order (optimization.h:85-100).
template <int KT, class PR>
__device__ __forceinline__ void shell_terms_reduced_k(const PR &pr, uint32_t s,
                                                      const uint32_t *cc, double &obj) {
  constexpr int L = KT - 1, NT = KT * (KT + 1) / 2;
  uint32_t c[NT];
  uint32_t fl = 0;  // sum over a < L of F[a][L]
  int t = 0;
#pragma unroll
  for (int xi = 0; xi < KT; ++xi)
#pragma unroll
    for (int eta = xi; eta < KT; ++eta, ++t) {
      if (eta < L) {
        c[t] = cc[xi * L + eta] + (xi == eta ? 0u : cc[eta * L + xi]);
      } else if (xi < L) {
        uint32_t f = 0;
#pragma unroll
        for (int b = 0; b < L; ++b) f += cc[xi * L + b] + cc[b * L + xi];
        c[t] = constant_ref(pr.dsum, s * KT + xi) - f;
        fl += c[t];
      } else {
        c[t] = (constant_ref(pr.dsum, s * KT + L) - fl) >> 1;
      }
    }
  const TermParam *tp = pr.tparams + s * NT;
#pragma unroll
  for (int u = 0; u < NT; ++u) {
    const auto &e = constant_ref(tp, u);
    const double v = 1.0 - static_cast<double>(c[u]) * e.pref;
    obj += e.pw * term_abs(v - e.tgt);
  }
}

template <class PR>
__device__ __forceinline__ void shell_terms_reduced(const PR &pr, uint32_t s,
                                                    const uint32_t *cc, double &obj) {
  switch (pr.K) {  // the host reduces problems of 2-6 species only
    case 2: return shell_terms_reduced_k<2>(pr, s, cc, obj);
    case 3: return shell_terms_reduced_k<3>(pr, s, cc, obj);
    case 4: return shell_terms_reduced_k<4>(pr, s, cc, obj);
    case 5: return shell_terms_reduced_k<5>(pr, s, cc, obj);
    default: return shell_terms_reduced_k<6>(pr, s, cc, obj);
  }
}

// Scores the candidates in LDS rows [0, npass * CW) (row c at rows + c * stride, 32 * dT
// bytes, padding sites 0xff) with Kc one-hot columns per candidate: pr.K, or pr.K - 1 where
// the host allows the last species to be left out (pr.Kc; search kernels only, no bond or
// SRO output), which fits 32 / (K - 1) candidates in a tile group instead of 32 / K.  `cnt`
// is CW * (Kc * Kc | 1) + 32 words of LDS.  Passes whose bit in `live` are clear are skipped.
// With `prune_key` below ~0, a pass stops after the first shell at which no active candidate
// of it (bit set in `active`) has a partial objective key at or below prune_key: every term
// is >= 0 (the host checks the pair weights), so partial sums only grow and such candidates
// can never be kept.  Returns the objective of the lane's own candidate (lane < npass * CW)
// and, when given, writes its ordered bonds and SRO cube.
template <int TT, int G, int D, class PR>
__device__ __forceinline__ double dense_score(const PR &pr, uint32_t Kc, const uint8_t *rows,
                                              uint32_t stride, uint32_t *cnt,
                                              uint32_t CW, uint32_t npass,
                                              unsigned long long live, int lane, int32_t *bonds,
                                              double *sro, uint32_t out_rows,
                                              unsigned long long active = 0,
                                              unsigned long long prune_key = ~0ull) {
  constexpr int P = TT * (TT + 1) / 2;
  const uint32_t K = Kc, KK = K * K, T = pr.dT, S = pr.S;
  const bool reduced = Kc != pr.K;
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
  const uint32_t cube = pr.S * pr.K * pr.K;
  // a candidate's counts take an odd number of words, so the lanes that read different
  // candidates' counts together hit different LDS banks
  const uint32_t cs = KK | 1u;
  // a spare word past the counts (lanes l and l + 32 share one: its value is never read)
  const uint32_t trash = CW * cs + (static_cast<uint32_t>(lane) & 31u);
  // count word of accumulator row m of tile group gi: candidate gi * g + cl, row m - cl * K
  const uint32_t cbase = cl * cs - cl * K * K + b;
  // accumulator registers of GEMM2 on this lane's diagonal block
  const uint32_t keep = diag_registers(cl * K, K, h, col_ok);
  double obj = 0.0;
  // a shell's terms (optimization.h:85-100), added by each candidate's own lane in the
  // reference order
  auto terms_of = [&](uint32_t opass, uint32_t os) {
    if (my_pass == opass) {
      if (reduced) {
        shell_terms_reduced(pr, os, cnt + my_cc * cs, obj);
      } else {
        const uint32_t cr = static_cast<uint32_t>(lane);  // candidate row
        const bool out = cr < out_rows;
        shell_terms(pr, os, cnt + my_cc * cs, obj, out && bonds ? bonds + cr * cube : nullptr,
                    out && sro ? sro + cr * cube : nullptr);
      }
    }
  };
  for (uint32_t pass = 0; pass < npass; ++pass) {
    if (!((live >> pass) & 1ull)) continue;
    // batched row reads up to the 8-tile bound (+0.8% at 256 sites); the larger bounds'
    // registers do not hold them
    v4i X[G][TT];
    onehot_tiles<TT, G, (TT <= 8)>(X, rows, pass * CW, g, cl, col_ok, stride, h, T, bb);
    // ring slot q % D holds grid position q (row-major over I <= J < TT) of the shell.
    // Grid positions are stepped incrementally so that every index folds to a constant
    // once the loops are unrolled.
    v4i ring[D];
#pragma unroll
    for (int q = 0; q < D; ++q) ring[q] = tile_load(rs, toff, tile_offset<TT>(0, q));
    for (uint32_t s = 0; s < S; ++s) {
      v16i C[G], Y[G];
#pragma unroll
      for (int gi = 0; gi < G; ++gi) C[gi] = v16i{};
      int I = 0, J = 0;  // current grid position
#pragma unroll
      for (int q = 0; q < P; ++q) {
        if (J == I) {
#pragma unroll
          for (int gi = 0; gi < G; ++gi) Y[gi] = v16i{};
        }
        const v4i u = ring[q % D];
        // the position D ahead, in the next shell past this one's end
        ring[q % D] = q + D < P ? tile_load(rs, toff, tile_offset<TT>(s, q + D))
                                : tile_load(rs, toff, tile_offset<TT>(s + 1, q + D - P));
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
      // row m = 8(r/4) + 4h + r%4 = cK + a of column n = cK + b.  Every register is
      // stored, the ones outside the lane's diagonal block into a spare word past the
      // counts: a select per store instead of a branch
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
#pragma unroll
        for (uint32_t r = 0; r < 16; ++r) {
          const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
          const uint32_t at = ((keep >> r) & 1u) ? gi * g * cs + cbase + K * m : trash;
          cnt[at] = static_cast<uint32_t>(C[gi][r]);
        }
      wave_sync();
      terms_of(pass, s);
      wave_sync();  // the counts are rewritten by the next shell
      if (prune_key != ~0ull && s + 1 < S) {
        const bool alive = my_pass == pass && ((active >> lane) & 1ull) &&
                           okey(obj) <= prune_key;
        if (__ballot(alive) == 0) break;
      }
    }
  }
  return obj;
}

// the sorted configuration (32 * dT bytes, dT <= TT) into an LDS row, 16 bytes at a time
template <int TT, class PR>
__device__ __forceinline__ void load_dense_row(uint8_t *row, const PR &pr) {
  const uint4 *src = reinterpret_cast<const uint4 *>(pr.species0D);
  uint4 *dst = reinterpret_cast<uint4 *>(row);
#pragma unroll
  for (uint32_t w = 0; w < 2 * TT; ++w)
    if (w < 2 * pr.dT) dst[w] = src[w];
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
  const bool gl = static_cast<uint32_t>(lane) < GEN;
  uint8_t *row = rows + (gl ? static_cast<uint32_t>(lane) : 0u) * stride;
  for (uint32_t r = 0; r < p.rounds; ++r) {
    const uint64_t cand = p.systematic ? sbeg + static_cast<uint64_t>(lane) * p.rounds + r
                                       : sbeg + static_cast<uint64_t>(r) * GEN + lane;
    const bool active = gl && cand < end;
    if (gl) {
      if (p.systematic) {
        if (r == 0) {
          load_dense_row<TT>(row, pr);
          if (active) unrank_candidate(Strided<0>{row}, pr, cand + 1, 1u);
        } else if (active) {
          next_perm_candidate(Strided<0>{row}, pr.bounds[0], pr.bounds[1] - pr.bounds[0]);
        }
      } else {
        load_dense_row<TT>(row, pr);
        if (active && !(p.chunk_quirk && block_start(p, cand)))
          shuffle_candidate(Strided<0>{row}, pr, cand);
      }
    }
    const unsigned long long am = __ballot(active);
    if (am == 0) break;  // the tail of a systematic segment
    unsigned long long live = 0;
    for (uint32_t q = 0; q < npass; ++q)
      if ((am >> (q * CW)) & low_mask(CW)) live |= 1ull << q;
    wave_sync();
    // early exit bound: the job threshold and the wavefront filter's bound (neither changes
    // during the round, so the record test below reads the same values after scoring)
    unsigned long long prune_key = ~0ull;
    if (p.prune) {
      const unsigned long long g0 =
          p.thr_key;
      const unsigned long long l0 = filter_bound(lf, p);
      prune_key = g0 < l0 ? g0 : l0;
    }
    const double tot =
        dense_score<TT, G, D>(pr, pr.Kc, rows, stride, cnt, CW, npass, live, lane, nullptr,
                              nullptr, 0u, am, prune_key);
    scored += static_cast<unsigned long long>(__popcll(am));
    // a candidate above the threshold, or above the filter's L-th level when its turn
    // comes, cannot be inserted
    const unsigned long long g =
        p.thr_key;
    const bool below = active && okey(tot) <= g;
    const bool pass = (filter_pass(lf, p, __ballot(below), tot) >> lane) & 1ull;
    uint32_t slot = ~0u;
    if (pass) {
      slot = atomicAdd(p.rec_count, 1u);
      if (slot < p.rec_cap) {
        p.rec_cand[slot] = cand;
        p.rec_obj[slot] = tot;
        p.rec_sub[slot] = tot;
      }
    }
    record_rows(p.rec_species, p.rec_stride, p.rec_cap, __ballot(pass), slot, rows, stride, pr.N,
                0u, lane);
    wave_sync();  // rows are rewritten by the next round of candidates
  }
}

// One segment of a split-mode job (sqs.h:404-418): a candidate holds one configuration per
// sublattice problem, each shuffled with its own stream (split mode shuffles every candidate,
// sqs.h:407); the problems share the kernel's tile bound and the pass width of the largest
// species count, and are scored one after the other.  The objective is 0.0 plus the
// problems' objectives in order (helpers::sum folds from T(0), core/helpers/fold.h:51-53);
// each lane's per-problem objectives wait in LDS (off_obj, [problem][lane]).
template <int TT, int G, int D, class RP>
__device__ __forceinline__ void dense_segment_split(const RP &p, uint64_t sbeg, uint64_t end,
                                                    uint8_t *wl, int lane, LevelFilter &lf,
                                                    unsigned long long &scored) {
  const uint32_t GEN = p.gen_group, CW = p.dense_cw, stride = p.dense_stride;
  const uint32_t npass = GEN / CW, nprob = p.nprob;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(wl + p.off_hist);
  double *subs = reinterpret_cast<double *>(wl + p.off_obj);
  const bool gl = static_cast<uint32_t>(lane) < GEN;
  const uint64_t cand = sbeg + static_cast<uint64_t>(lane);  // Monte Carlo: one round
  const bool active = gl && cand < end;
  for (uint32_t q = 0; q < nprob; ++q) {
    uint8_t *row = wl + p.off_rows[q] + (gl ? static_cast<uint32_t>(lane) : 0u) * stride;
    if (gl) {
      load_dense_row<TT>(row, p.prob[q]);
      if (active) shuffle_candidate(Strided<0>{row}, p.prob[q], cand);
    }
  }
  const unsigned long long am = __ballot(active);
  if (am == 0) return;
  unsigned long long live = 0;
  for (uint32_t q = 0; q < npass; ++q)
    if ((am >> (q * CW)) & low_mask(CW)) live |= 1ull << q;
  wave_sync();
  double tot = 0.0;
  for (uint32_t q = 0; q < nprob; ++q) {
    const double o = dense_score<TT, G, D>(
        p.prob[q], p.prob[q].Kc, wl + p.off_rows[q], stride, cnt, CW, npass, live, lane, nullptr,
        nullptr, 0u, am, ~0ull);
    subs[q * 64 + static_cast<uint32_t>(lane)] = o;
    tot += o;
  }
  scored += static_cast<unsigned long long>(__popcll(am));
  const unsigned long long g = p.thr_key;
  const bool below = active && okey(tot) <= g;
  const bool pass = (filter_pass(lf, p, __ballot(below), tot) >> lane) & 1ull;
  uint32_t slot = ~0u;
  if (pass) {
    slot = atomicAdd(p.rec_count, 1u);
    if (slot < p.rec_cap) {
      for (uint32_t q = 0; q < nprob; ++q)
        p.rec_sub[slot * nprob + q] = subs[q * 64 + static_cast<uint32_t>(lane)];
      p.rec_cand[slot] = cand;
      p.rec_obj[slot] = tot;
    }
  }
  const unsigned long long pm = __ballot(pass);
  uint32_t off = 0;
  for (uint32_t q = 0; q < nprob; ++q) {
    record_rows(p.rec_species, p.rec_stride, p.rec_cap, pm, slot, wl + p.off_rows[q], stride,
                p.prob[q].N, off, lane);
    off += p.prob[q].N;
  }
  wave_sync();  // rows are rewritten by the next segment
}

// The two wavefronts of a SIMD start together and, doing the same work at the same pace, stay
// in step: both shuffle candidates (vector work only) at the same time, then both contend
// for the matrix core.  The second half of the grid (blocks from `from`, half the grid as
// the host passes it), which the dispatcher places on the SIMDs beside the first half,
// starts about 16k cycles later, so that one wavefront's vector phases tend to fall in the
// other's MFMA phases (+1-2% at 256 sites; at worst a fraction of a percent of a launch
// where the placement differs).  The grid size comes from the parameters: reading it from
// the dispatch packet holds two more scalar registers through the kernel, which spilled.
__device__ __forceinline__ void stagger_start(uint32_t from) {
  if (blockIdx.x >= from) {
    __builtin_amdgcn_s_sleep(127);
    __builtin_amdgcn_s_sleep(127);
  }
}

template <int TT, int G, int W, int D, bool SPLIT>
__global__ __launch_bounds__(64, W) void run_dense_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  stagger_start(p.stagger_from);
  const uint64_t per_seg = static_cast<uint64_t>(p.gen_group) * p.rounds;
  const uint64_t total = p.cand_end - p.cand_begin;
  const uint64_t nseg = (total + per_seg - 1) / per_seg;
  // the tie filter's levels are the same on every lane: one copy per wavefront in LDS
  LevelFilter &lf = *reinterpret_cast<LevelFilter *>(lds + p.off_levels);
  lf.n = 0;
  unsigned long long scored = 0;
  // segments are pulled from a counter: wavefronts that start late (more blocks than fit
  // at once) or run slow do not hold up the launch
  for (;;) {
    uint32_t seg = 0;
    if (lane == 0) seg = atomicAdd(p.next_seg, 1u);
    seg = __builtin_amdgcn_readfirstlane(__shfl(seg, 0));
    if (seg >= nseg) break;
    if constexpr (SPLIT)
      dense_segment_split<TT, G, D>(p, p.cand_begin + seg * per_seg, p.cand_end, lds, lane, lf,
                                    scored);
    else
      dense_segment<TT, G, D>(p, p.cand_begin + seg * per_seg, p.cand_end, lds, lane, lf,
                              scored);
  }
  if (lane == 0 && scored) atomicAdd(p.scored, scored);
}

// Persistent form: every wavefront claims one-segment chunks of the jobs' candidate ranges
// until the work list is drained; the jobs of a launch share the tile bound TT.
template <int TT, int G, int W, int D, bool SPLIT>
__global__ __launch_bounds__(64, W) void batch_dense_kernel(BatchParams bp) {
  extern __shared__ __align__(16) uint8_t lds[];
  const int lane = threadIdx.x;
  stagger_start(bp.stagger_from);
  const uint32_t num_items = *bp.num_items;
  uint32_t last_job = 0xffffffffu, it = 0, job = 0;
  uint64_t cbeg = 0, cend = 0;
  LevelFilter *lf = nullptr;  // the job's tie filter levels, one copy per wavefront in LDS
  unsigned long long scored = 0;
  while (claim_chunk(bp, num_items, it, lane, job, cbeg, cend)) {
    // the job table is read-only during the launch: scalar loads through the constant
    // address space keep every job parameter uniform
    const auto &p = constant_ref(bp.jobs, job);
    if (job != last_job) {
      if (last_job != 0xffffffffu && lane == 0 && scored)
        atomicAdd(bp.jobs[last_job].scored, scored);
      scored = 0;
      lf = reinterpret_cast<LevelFilter *>(lds + p.off_levels);
      lf->n = 0;
      last_job = job;
    }
    if constexpr (SPLIT)
      dense_segment_split<TT, G, D>(p, cbeg, cend, lds, lane, *lf, scored);
    else
      dense_segment<TT, G, D>(p, cbeg, cend, lds, lane, *lf, scored);
  }
  if (last_job != 0xffffffffu && lane == 0 && scored) atomicAdd(bp.jobs[last_job].scored, scored);
}

// LDS rows of explicit candidates [b0, b0 + n) of a B x N batch (row stride `stride`,
// 32 T bytes each, sites at or past N set to the padding byte 0xff); candidates at or past B
// get species 0 (scored, never reported).  The wavefront copies the rows together: with
// N a multiple of 16, 16-byte words of consecutive lanes read consecutive batch bytes.
__device__ __forceinline__ void stage_rows(uint8_t *rows, uint32_t stride, const uint8_t *species,
                                           uint64_t b0, uint32_t n, uint64_t B, uint32_t N,
                                           uint32_t T, int lane) {
  if ((N & 15u) == 0) {
    const uint32_t per_row = 2 * T, chunks = n * per_row;  // 16-byte words
#pragma unroll 4
    for (uint32_t w = static_cast<uint32_t>(lane); w < chunks; w += 64) {
      const uint32_t c = w / per_row, k = w - c * per_row;
      uint4 v = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu};
      if (16 * k < N)
        v = b0 + c < B ? *reinterpret_cast<const uint4 *>(species + (b0 + c) * N + 16 * k)
                       : uint4{0, 0, 0, 0};
      *reinterpret_cast<uint4 *>(rows + c * stride + 16 * k) = v;
    }
    return;
  }
  if (static_cast<uint32_t>(lane) >= n) return;  // one lane per row, byte by byte
  const uint64_t b = b0 + static_cast<uint32_t>(lane);
  uint8_t *row = rows + lane * stride;
  for (uint32_t i = 0; i < 32 * T; ++i)
    row[i] = (b < B && i < N) ? species[b * N + i] : (i < N ? 0 : 0xff);
}

template <int TT, int G, int D>
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
  stage_rows(rows, p.dense_stride, p.species, static_cast<uint64_t>(blockIdx.x) * GEN, GEN, p.B,
             pr.N, pr.dT, lane);
  const unsigned long long am = __ballot(active);
  unsigned long long live = 0;
  for (uint32_t q = 0; q < npass; ++q)
    if ((am >> (q * CW)) & low_mask(CW)) live |= 1ull << q;
  wave_sync();
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * GEN;
  const uint32_t out_rows = static_cast<uint32_t>(p.B - b0 < GEN ? p.B - b0 : GEN);
  const double obj = dense_score<TT, G, D>(
      pr, pr.K, rows, p.dense_stride, cnt, CW, npass, live, lane,
      p.bonds ? p.bonds + b0 * nb : nullptr, p.sro ? p.sro + b0 * nb : nullptr, out_rows);
  if (active) p.objective[b] = obj;
}

// ---- explicit scoring split by shell -----------------------------------------------------
//
// For batches too small to fill the device with whole candidates: a block of kShellWaves
// wavefronts scores one shell s (blockIdx.y) for kShellWaves passes of CW candidates (one
// pass per wavefront).  The block's wavefronts share the shell's tiles in LDS: LT (tile
// bounds up to 8, a shell's tiles at most 36 KiB) copies them once; larger bounds stream them
// through kShellBufs LDS buffers of shell_chunk() grid positions: the loads of chunk k + 3
// are issued as chunk k starts (two chunks in registers at a time) and chunk k + 2 is stored
// as chunk k ends, then one barrier.  Read from LDS, a tile costs 1 KiB of LDS bandwidth per
// MFMA group; read by each wavefront from the L2 stream it would cost the same in L1
// bandwidth, which holds half of LDS's.  The tile, term-parameter and row loads of a block
// are all issued before the first is waited for.  After the GEMMs, all 64 lanes of a
// wavefront work on its pass's outputs: the term values (lane items (term, candidate),
// candidate fastest) go to `vals` [T][B] (term-major: the sum kernel reads them coalesced),
// the SRO entries and the bond counts to the outputs.  score_sum_kernel then adds each
// candidate's T values from 0.0 in the reference order (optimization.h:85-100): the same
// values in the same order as the whole-candidate kernels, so objectives are bit-identical.
// Measured alternatives at 256 sites: folding the sum into the last of a column's S blocks
// to finish needs a release fence at agent scope (an L2 write-back) in every block, 2.3x
// slower; 16-bit bond counts instead of the values move the term arithmetic into the sum
// kernel's serial per-candidate loop, 1.4x slower.
constexpr uint32_t kShellWaves = 4;
constexpr int kShellBufs = 3;
// streamed form: one tile group per wavefront fits three blocks per CU in registers and, with
// chunks of 8 positions, in LDS (the 1197 blocks of 864 sites x 4096 candidates then run in
// two rounds instead of three); two groups keep chunks of 16 and one block per SIMD pair
template <int TT, int G> constexpr bool shell_three_blocks() { return G == 1 && TT <= 27; }
template <int TT, int G> constexpr int shell_chunk() { return shell_three_blocks<TT, G>() ? 8 : 16; }

template <int TT, int G, int D, bool LT>
__global__ __launch_bounds__(64 * kShellWaves, LT ? 2 : (shell_three_blocks<TT, G>() ? 3 : 1))
void score_shell_kernel(ScoreParams p, double *vals) {
  extern __shared__ __align__(16) uint8_t lds[];
  constexpr int W = static_cast<int>(kShellWaves);
  constexpr int P = TT * (TT + 1) / 2;
  constexpr int Q = shell_chunk<TT, G>(), NCK = (P + Q - 1) / Q;  // streamed: chunks of Q tiles
  constexpr int PER = Q * 64 / (64 * W);  // 16-byte loads of a chunk per thread
  static_assert(LT || (Q * 64) % (64 * W) == 0, "a chunk splits evenly over the block");
  constexpr uint32_t kTileLds = static_cast<uint32_t>(LT ? P : kShellBufs * Q) * 1024u;
  const uint32_t wave = threadIdx.x >> 6;
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const DevProblem &pr = p.prob;
  const uint32_t K = pr.K, KK = K * K, S = pr.S, T = pr.dT, s = blockIdx.y;
  const uint32_t NT = K * (K + 1) / 2;
  const uint32_t CW = p.dense_cw, g = CW / G, stride = p.dense_stride;
  // LT: the shell's P tiles (64 16-byte words each); else kShellBufs chunks of Q tiles
  v4i *tbuf = reinterpret_cast<v4i *>(lds);
  TermParam *tpar = reinterpret_cast<TermParam *>(lds + kTileLds);  // the shell's NT terms
  // the wavefront's candidate rows; after its one-hot tiles are built, its CW x K x K counts
  // (LDS operations of one wavefront stay in program order)
  uint8_t *rows = lds + p.off_row + wave * p.shell_rows;
  uint32_t *cnt = reinterpret_cast<uint32_t *>(rows);
  // this wavefront's pass: candidates [b0, b0 + CW)
  const uint64_t pass = static_cast<uint64_t>(blockIdx.x) * W + wave;
  const uint64_t b0 = pass * CW;
  // the shell's tiles in the stream.  Streamed chunks may run past the shell's last
  // position into the next shell (the stream ends with a zero shell), never past the stream.
  const v4i *stream = reinterpret_cast<const v4i *>(pr.dtiles) + tile_offset<TT>(s, 0) / 16;
  const int tid = static_cast<int>(threadIdx.x);
  // loads in flight: the shell's tiles (LT) or its first two chunks and the third in its
  // register set (chunk c in pf[c & 1]), the term parameters, then the rows
  constexpr int NCH = LT ? (P * 64 + 64 * W - 1) / (64 * W) : 2 * PER;
  v4i tl_in[NCH];
  v4i pf[LT ? 1 : 2][PER];
  if constexpr (LT) {
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int idx = tid + k * 64 * W;
      if (idx < P * 64) tl_in[k] = stream[idx];
    }
  } else {
#pragma unroll
    for (int k = 0; k < NCH; ++k)
      if (k < PER || NCK > 1) tl_in[k] = stream[tid + k * 64 * W];
    if (NCK > 2) {
#pragma unroll
      for (int j = 0; j < PER; ++j) pf[0][j] = stream[2 * Q * 64 + tid + j * 64 * W];
    }
  }
  const uint32_t nprm = NT * (sizeof(TermParam) / 16);  // 16-byte words of the parameters
  const uint4 *prm_src = reinterpret_cast<const uint4 *>(pr.tparams + s * NT);
  uint4 prm = {0, 0, 0, 0};
  if (threadIdx.x < nprm) prm = prm_src[threadIdx.x];
  stage_rows(rows, stride, p.species, b0, CW, p.B, pr.N, T, lane);
  if (threadIdx.x < nprm) reinterpret_cast<uint4 *>(tpar)[threadIdx.x] = prm;
  for (uint32_t i = threadIdx.x + 64 * W; i < nprm; i += 64 * W)  // K > 8 only
    reinterpret_cast<uint4 *>(tpar)[i] = prm_src[i];
  if constexpr (LT) {
#pragma unroll
    for (int k = 0; k < NCH; ++k) {
      const int idx = tid + k * 64 * W;
      if (idx < P * 64) tbuf[idx] = tl_in[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < NCH; ++k)
      if (k < PER || NCK > 1) tbuf[tid + k * 64 * W] = tl_in[k];
  }
  __syncthreads();  // tiles, parameters and rows in LDS
  const uint32_t h = static_cast<uint32_t>(lane) >> 5, n = static_cast<uint32_t>(lane) & 31u;
  const uint32_t cl = n / K, bcol = n - cl * K;
  const bool col_ok = cl < g;
  const uint32_t bb = bcol * 0x01010101u;
  // the long rows of the large bounds: every row read is issued before the first one-hot
  // conversion (-3% at 864 sites against one conditional read at a time)
  v4i X[G][TT];
  onehot_tiles<TT, G, !LT>(X, rows, 0u, g, cl, col_ok, stride, h, T, bb);
  v16i C[G], Y[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) C[gi] = v16i{};
  const v4i *tl = tbuf + lane;
  // grid positions q row-major over I <= J < TT, stepped so that indexes fold to constants
  int I = 0, J = 0;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (J == I) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) Y[gi] = v16i{};
    }
    const int k = q / Q, r = q % Q;  // streamed form: chunk k, position r in it
    if (!LT && r == 0 && k + 3 < NCK) {
#pragma unroll
      for (int j = 0; j < PER; ++j)
        pf[(k + 3) & 1][j] = stream[(k + 3) * Q * 64 + tid + j * 64 * W];
    }
    const v4i u = LT ? tl[q * 64] : tl[((k % kShellBufs) * Q + r) * 64];
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
      Y[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(u, X[gi][J], Y[gi], 0, 0, 0);
    if (J == TT - 1) {  // row I complete: C += X_I^T Y_I
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
        C[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[gi][I], pack_counts(Y[gi]), C[gi], 0,
                                                      0, 0);
    }
    if (++J == TT) ++I, J = I;
    // end of chunk k: chunk k + 2 into the buffer chunk k - 1 used (every wavefront is past
    // it: the barrier at the end of chunk k - 1), then the barrier that publishes it
    if (!LT && r == Q - 1 && k + 1 < NCK) {
      if (k + 2 < NCK) {
#pragma unroll
        for (int j = 0; j < PER; ++j)
          tbuf[((k + 2) % kShellBufs) * Q * 64 + tid + j * 64 * W] = pf[(k + 2) & 1][j];
      }
      __syncthreads();
    }
  }
  // diagonal blocks -> counts [candidate of the pass][a][b]: register r of lane l is row
  // m = 8(r/4) + 4h + r%4 = cK + a of column n = cK + b
  const uint32_t base = cl * K;
#pragma unroll
  for (int gi = 0; gi < G; ++gi)
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) {
      const uint32_t a = 8 * (r >> 2) + 4 * h + (r & 3) - base;  // wraps below the block
      if (col_ok && a < K) cnt[(gi * g + cl) * KK + a * K + bcol] = static_cast<uint32_t>(C[gi][r]);
    }
  wave_sync();
  // the pass's term values of shell s, its SRO entries and bond counts
  const uint32_t cube = S * KK, t0 = s * NT;
  for (uint32_t i = static_cast<uint32_t>(lane); i < NT * CW; i += 64) {
    const uint32_t t = i / CW, c = i - t * CW;
    const uint64_t b = b0 + c;
    if (b >= p.B) continue;
    const TermParam &e = tpar[t];
    const uint32_t *cc = cnt + c * KK;
    const uint32_t cab = cc[e.xi * K + e.eta];
    const uint32_t cba = e.xi == e.eta ? 0u : cc[e.eta * K + e.xi];
    const double v = 1.0 - static_cast<double>(cab + cba) * e.pref;
    vals[static_cast<uint64_t>(t0 + t) * p.B + b] = e.pw * absolute(v - e.tgt);
    if (p.sro) {
      double *o = p.sro + b * cube + s * KK;
      o[e.xi * K + e.eta] = v;
      o[e.eta * K + e.xi] = v;
    }
  }
  if (p.bonds) {
    for (uint32_t i = static_cast<uint32_t>(lane); i < CW * KK; i += 64) {
      const uint32_t c = i / KK, k = i - c * KK;
      const uint64_t b = b0 + c;
      if (b < p.B) p.bonds[b * cube + s * KK + k] = static_cast<int32_t>(cnt[i]);
    }
  }
}

// objectives of kSumCands candidates per block: their T term values, staged through LDS in
// chunks of kSumTerms terms (loads spread over the block, coalesced along candidates), added
// to 0.0 in the reference order by one thread per candidate
constexpr uint32_t kSumCands = 16, kSumTerms = 256, kSumThreads = 256;

__global__ __launch_bounds__(kSumThreads) void score_sum_kernel(ScoreParams p,
                                                               const double *vals) {
  __shared__ double v[kSumTerms * kSumCands];
  const uint64_t b0 = static_cast<uint64_t>(blockIdx.x) * kSumCands;
  const uint32_t T = p.prob.T;
  double obj = 0.0;
  for (uint32_t t0 = 0; t0 < T; t0 += kSumTerms) {
    constexpr uint32_t kPer = kSumTerms * kSumCands / kSumThreads;
    double x[kPer];
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
      const uint32_t i = threadIdx.x + k * kSumThreads;
      const uint32_t t = t0 + i / kSumCands, c = i % kSumCands;
      x[k] = t < T && b0 + c < p.B ? vals[static_cast<uint64_t>(t) * p.B + b0 + c] : 0.0;
    }
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) v[threadIdx.x + k * kSumThreads] = x[k];
    __syncthreads();
    if (threadIdx.x < kSumCands) {
      // eight values loaded together, then added one after another in order
      const uint32_t nt = T - t0 < kSumTerms ? T - t0 : kSumTerms;
      const double *w = v + threadIdx.x;
      uint32_t t = 0;
      for (; t + 8 <= nt; t += 8) {
        double x[8];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) x[k] = w[(t + k) * kSumCands];
#pragma unroll
        for (uint32_t k = 0; k < 8; ++k) obj += x[k];
      }
      for (; t < nt; ++t) obj += w[t * kSumCands];
    }
    __syncthreads();  // the next chunk overwrites the stage
  }
  if (threadIdx.x < kSumCands && b0 + threadIdx.x < p.B) p.objective[b0 + threadIdx.x] = obj;
}

// ---- paired form (MC jobs) -------------------------------------------------------------
//
// A block is one CU's worth of wavefronts: four pairs of an MFMA wavefront (M, waves 0-3)
// and a helper (H, waves 4-7; wave w + 4 shares the SIMD of wave w).  All eight advance
// on a common tick, one tick per (pass, shell) of a batch of gen_group candidates per pair:
//   M at a tick of batch b: the one-hot tiles (first shell of a pass), the shell's GEMMs
//     and its diagonal-block counts into cnt[tick parity];
//   H at the same tick: the objective terms of the previous tick's counts (summed per
//     candidate in reference order); at a batch's first tick the end of batch b - 1 (tie
//     filter, records); at every tick a slice of the Fisher-Yates shuffles of batch b + 1.
// Candidate rows are double-buffered by batch parity, counts by tick parity, and one
// __syncthreads per tick separates producers and consumers.  Wave 4 claims the block's
// batches two ahead into three LDS descriptors; every wave reads the same descriptors, so
// all run the same ticks and barriers.  Without generation and objective code the MFMA
// wavefront keeps its registers for the one-hot tiles, the tile ring and the accumulators,
// and the helper's vector and LDS work runs beside the matrix core on the same SIMD.
//
// LDS: descriptors in [0, 128), then pair p's region at 128 + p * ws_pair: rows
// off_rows[0] / off_rows[1], counts off_hist / off_terms.

constexpr uint32_t kWsPairs = kDensePairs, kWsHead = kDensePairHead;

struct WsBatch {  // a block batch: candidates [cbeg, cend) of job `job`
  uint64_t cbeg, cend;
  uint32_t job, valid;
};

// claim sources: the single-job kernel takes block batches of one job from a counter, the
// persistent kernel one-batch chunks of a work list
struct RunSource {
  const RunParams &p;
  __device__ WsBatch next(int lane) {
    const uint64_t per = static_cast<uint64_t>(kWsPairs) * p.gen_group;
    const uint64_t nseg = (p.cand_end - p.cand_begin + per - 1) / per;
    uint32_t seg = 0;
    if (lane == 0) seg = atomicAdd(p.next_seg, 1u);
    seg = __builtin_amdgcn_readfirstlane(__shfl(seg, 0));
    if (seg >= nseg) return WsBatch{0, 0, 0, 0};
    const uint64_t b = p.cand_begin + seg * per;
    return WsBatch{b, b + per < p.cand_end ? b + per : p.cand_end, 0, 1};
  }
  __device__ const RunParams &params(uint32_t) const { return p; }
};

struct BatchSource {
  const BatchParams &bp;
  uint32_t num_items, it;
  __device__ WsBatch next(int lane) {
    uint32_t job = 0;
    uint64_t cb = 0, ce = 0;
    if (!claim_chunk(bp, num_items, it, lane, job, cb, ce)) return WsBatch{0, 0, 0, 0};
    return WsBatch{cb, ce, job, 1};
  }
  __device__ const __attribute__((address_space(4))) RunParams &params(uint32_t job) const {
    return constant_ref(bp.jobs, job);
  }
};

// M: one (pass, shell).  X and the ring live across the ticks of a pass: the ring
// prefetches the next shell's first positions (zeros past the last shell) and is refilled
// with the X tiles at the first shell of a pass.
template <int TT, int G, int D, class PR>
__device__ __forceinline__ void ws_mfma_tick(const PR &pr, const uint8_t *rows, uint32_t stride,
                                             uint32_t CW, uint32_t pass, uint32_t s,
                                             v4i (&X)[G][TT], v4i (&ring)[D], uint32_t *cnt,
                                             int lane) {
  constexpr int P = TT * (TT + 1) / 2;
  const uint32_t K = pr.K, KK = K * K, T = pr.dT, S = pr.S;
  const uint32_t g = CW / G;
  const uint32_t h = static_cast<uint32_t>(lane) >> 5, n = static_cast<uint32_t>(lane) & 31u;
  const uint32_t cl = n / K, b = n - cl * K;
  const bool col_ok = cl < g;
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<uint32_t *>(pr.dtiles), static_cast<short>(0),
      static_cast<int>(pr.dtiles_n * 1024u), 0x00020000);
  const uint32_t toff = static_cast<uint32_t>(lane) * 16u;
  if (s == 0) {
    onehot_tiles<TT, G, false>(X, rows, pass * CW, g, cl, col_ok, stride, h, T,
                               b * 0x01010101u);
#pragma unroll
    for (int q = 0; q < D; ++q) ring[q] = tile_load(rs, toff, tile_offset<TT>(0, q));
  }
  v16i C[G], Y[G];
#pragma unroll
  for (int gi = 0; gi < G; ++gi) C[gi] = v16i{};
  int I = 0, J = 0;
#pragma unroll
  for (int q = 0; q < P; ++q) {
    if (J == I) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi) Y[gi] = v16i{};
    }
    const v4i u = ring[q % D];
    ring[q % D] = q + D < P ? tile_load(rs, toff, tile_offset<TT>(s, q + D))
                            : tile_load(rs, toff, tile_offset<TT>(s + 1, q + D - P));
#pragma unroll
    for (int gi = 0; gi < G; ++gi)
      Y[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(u, X[gi][J], Y[gi], 0, 0, 0);
    if (J == TT - 1) {
#pragma unroll
      for (int gi = 0; gi < G; ++gi)
        C[gi] = __builtin_amdgcn_mfma_i32_32x32x32_i8(X[gi][I], pack_counts(Y[gi]), C[gi], 0,
                                                      0, 0);
    }
    if (++J == TT) ++I, J = I;
  }
#pragma unroll
  for (int gi = 0; gi < G; ++gi)
#pragma unroll
    for (uint32_t r = 0; r < 16; ++r) {
      const uint32_t m = 8 * (r >> 2) + 4 * h + (r & 3);
      if (col_ok && m / K == cl) cnt[gi * g * KK + K * m + b] = static_cast<uint32_t>(C[gi][r]);
    }
}

// H: the objective terms of one (pass, shell) from its counts, added to the running
// objectives of the pass's candidates (the lane of candidate row c holds its sum)
template <class PR>
__device__ __forceinline__ void ws_objective(const PR &pr, uint32_t CW, uint32_t pass,
                                             uint32_t s, const uint32_t *cnt, double &obj,
                                             int lane) {
  const uint32_t c = static_cast<uint32_t>(lane) - pass * CW;
  if (c < CW) shell_terms(pr, s, cnt + c * pr.K * pr.K, obj, nullptr, nullptr);
}

// H: Fisher-Yates state of the lane's candidate (the steps of shuffle_candidate spread
// over ticks): window w, its low bound and the current i
struct FyState {
  uint64_t st;
  uint32_t w, lo, i;
};

template <class PR>
__device__ __forceinline__ void fy_init(FyState &f, const PR &pr, uint64_t cand, bool shuffle) {
  f.st = pr.seed + (cand * static_cast<uint64_t>(pr.draws)) * kSecret0;
  f.w = shuffle ? 0u : pr.nbounds;
  f.lo = pr.bounds[0];
  f.i = pr.bounds[1] - pr.bounds[0];
}

// at most `n` swaps
template <class PR>
__device__ __forceinline__ void fy_steps(FyState &f, const PR &pr, uint8_t *row, uint32_t n) {
  while (n > 0 && f.w < pr.nbounds) {
    if (f.i > 1) {
      const uint32_t q = bounded(draw32(f.st), f.i);
      const uint8_t tmp = row[f.lo + f.i - 1];
      row[f.lo + f.i - 1] = row[f.lo + q];
      row[f.lo + q] = tmp;
      --f.i;
      --n;
    } else if (++f.w < pr.nbounds) {
      f.lo = pr.bounds[2 * f.w];
      f.i = pr.bounds[2 * f.w + 1] - f.lo;
    }
  }
}

// H: tie filter and records of a pair's finished batch [cbeg, cend) (as in dense_segment)
template <class RP>
__device__ __forceinline__ void ws_record(const RP &p, const uint8_t *rows, uint64_t cbeg,
                                          uint64_t cend, double tot, LevelFilter &lf,
                                          int lane) {
  const auto &pr = p.prob[0];
  const uint64_t cand = cbeg + static_cast<uint32_t>(lane);
  const bool active = static_cast<uint32_t>(lane) < p.gen_group && cand < cend;
  const unsigned long long am = __ballot(active);
  if (am == 0) return;
  if (lane == 0) atomicAdd(p.scored, static_cast<unsigned long long>(__popcll(am)));
  const unsigned long long g =
      p.thr_key;
  const bool below = active && okey(tot) <= g;
  const bool pass = (filter_pass(lf, p, __ballot(below), tot) >> lane) & 1ull;
  uint32_t slot = ~0u;
  if (pass) {
    slot = atomicAdd(p.rec_count, 1u);
    if (slot < p.rec_cap) {
      p.rec_cand[slot] = cand;
      p.rec_obj[slot] = tot;
      p.rec_sub[slot] = tot;
    }
  }
  record_rows(p.rec_species, p.rec_stride, p.rec_cap, __ballot(pass), slot, rows, p.dense_stride,
              pr.N, 0u, lane);
}

// H: starts the pair's share of block batch `bt` in `rows` (sorted rows, shuffle state)
template <int TT, class RP>
__device__ __forceinline__ void ws_start(const RP &p, const WsBatch &bt, uint32_t pair,
                                         uint8_t *rows, FyState &f, int lane) {
  if (static_cast<uint32_t>(lane) >= p.gen_group) return;
  const auto &pr = p.prob[0];
  load_dense_row<TT>(rows + static_cast<uint32_t>(lane) * p.dense_stride, pr);
  const uint64_t cand = bt.cbeg + pair * p.gen_group + static_cast<uint32_t>(lane);
  fy_init(f, pr, cand, cand < bt.cend && !(p.chunk_quirk && block_start(p, cand)));
}

template <int TT, int G, int D, class Source>
__device__ __forceinline__ void ws_body(Source &src, uint8_t *lds) {
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6;
  WsBatch *info = reinterpret_cast<WsBatch *>(lds);
  if (wave < kWsPairs) {  // ---- M ----
    v4i X[G][TT];
    v4i ring[D];
    __syncthreads();  // batches 0 and 1 claimed
    __syncthreads();  // batch 0 generated
    uint32_t tick_no = 0;
    for (uint32_t b = 0;; ++b) {
      const WsBatch cur = info[b % 3];
      if (!cur.valid) break;
      const auto &p = src.params(cur.job);
      const auto &pr = p.prob[0];
      const uint32_t S = pr.S, CW = p.dense_cw, n = (p.gen_group / CW) * S;
      uint8_t *base = lds + kWsHead + wave * p.ws_pair;
      const uint8_t *rows = base + p.off_rows[b & 1];
      uint32_t pass = 0, s = 0;
      for (uint32_t tick = 0; tick < n; ++tick, ++tick_no) {
        uint32_t *cnt = reinterpret_cast<uint32_t *>(base + ((tick_no & 1) ? p.off_terms
                                                                           : p.off_hist));
        ws_mfma_tick<TT, G, D>(pr, rows, p.dense_stride, CW, pass, s, X, ring, cnt, lane);
        if (++s == S) s = 0, ++pass;
        __syncthreads();
      }
    }
  } else {  // ---- H ----
    const uint32_t pair = wave - kWsPairs;
    LevelFilter lf;
    lf.n = 0;
    uint32_t lf_job = 0xffffffffu;
    FyState f{};
    double obj = 0.0;
    if (wave == kWsPairs) {
      const WsBatch b0 = src.next(lane);
      const WsBatch b1 = b0.valid ? src.next(lane) : WsBatch{0, 0, 0, 0};
      if (lane == 0) info[0] = b0, info[1] = b1, info[2] = WsBatch{0, 0, 0, 0};
    }
    __syncthreads();
    {
      const WsBatch b0 = info[0];
      if (b0.valid) {
        const auto &p = src.params(b0.job);
        uint8_t *rows = lds + kWsHead + pair * p.ws_pair + p.off_rows[0];
        ws_start<TT>(p, b0, pair, rows, f, lane);
        if (static_cast<uint32_t>(lane) < p.gen_group)
          fy_steps(f, p.prob[0], rows + static_cast<uint32_t>(lane) * p.dense_stride,
                   0xffffffffu);
      }
    }
    wave_sync();
    __syncthreads();
    WsBatch prev{0, 0, 0, 0};
    uint32_t tick_no = 0, prev_tick = 0;
    // the last shell, tie filter and records of batch b - 1 (prev)
    auto finish_prev = [&](uint32_t b) {
      const auto &pp = src.params(prev.job);
      uint8_t *pbase = lds + kWsHead + pair * pp.ws_pair;
      const uint32_t pCW = pp.dense_cw;
      ws_objective(pp.prob[0], pCW, pp.gen_group / pCW - 1, pp.prob[0].S - 1,
                   reinterpret_cast<const uint32_t *>(
                       pbase + ((prev_tick & 1) ? pp.off_terms : pp.off_hist)),
                   obj, lane);
      if (prev.job != lf_job) lf.n = 0, lf_job = prev.job;
      ws_record(pp, pbase + pp.off_rows[(b - 1) & 1], prev.cbeg + pair * pp.gen_group,
                prev.cend, obj, lf, lane);
    };
    for (uint32_t b = 0;; ++b) {
      const WsBatch cur = info[b % 3];
      if (!cur.valid) {
        if (prev.valid) finish_prev(b);
        break;
      }
      const WsBatch nxt = info[(b + 1) % 3];
      const auto &p = src.params(cur.job);
      const auto &pr = p.prob[0];
      const uint32_t S = pr.S, CW = p.dense_cw, n = (p.gen_group / CW) * S;
      uint8_t *base = lds + kWsHead + pair * p.ws_pair;
      uint32_t steps = 0;
      uint32_t pass = 0, s = 0;  // (pass, shell) of the previous tick of this batch
      for (uint32_t tick = 0; tick < n; ++tick, ++tick_no) {
        if (tick == 0) {
          if (prev.valid) finish_prev(b);
          obj = 0.0;
          if (wave == kWsPairs) {  // claims stay two batches ahead
            const WsBatch b2 = nxt.valid ? src.next(lane) : WsBatch{0, 0, 0, 0};
            if (lane == 0) info[(b + 2) % 3] = b2;
          }
          if (nxt.valid) {
            const auto &np = src.params(nxt.job);
            ws_start<TT>(np, nxt, pair,
                     lds + kWsHead + pair * np.ws_pair + np.off_rows[(b + 1) & 1], f, lane);
            steps = (np.prob[0].draws + n - 1) / n;
          }
        } else {
          ws_objective(pr, CW, pass, s,
                       reinterpret_cast<const uint32_t *>(
                           base + (((tick_no - 1) & 1) ? p.off_terms : p.off_hist)),
                       obj, lane);
          if (++s == S) s = 0, ++pass;
        }
        if (nxt.valid) {  // a slice of the next batch's shuffles, the rest at the last tick
          const auto &np = src.params(nxt.job);
          if (static_cast<uint32_t>(lane) < np.gen_group)
            fy_steps(f, np.prob[0],
                     lds + kWsHead + pair * np.ws_pair + np.off_rows[(b + 1) & 1] +
                         static_cast<uint32_t>(lane) * np.dense_stride,
                     tick + 1 == n ? 0xffffffffu : steps);
        }
        wave_sync();
        __syncthreads();
      }
      prev = cur;
      prev_tick = tick_no - 1;
    }
  }
}

template <int TT, int G, int D>
__global__ __launch_bounds__(512, 2) void run_dense_ws_kernel(RunParams p) {
  extern __shared__ __align__(16) uint8_t lds[];
  RunSource src{p};
  ws_body<TT, G, D>(src, lds);
}

template <int TT, int G, int D>
__global__ __launch_bounds__(512, 2) void batch_dense_ws_kernel(BatchParams bp) {
  extern __shared__ __align__(16) uint8_t lds[];
  BatchSource src{bp, *bp.num_items, 0u};
  ws_body<TT, G, D>(src, lds);
}

hipError_t dense_lds(const void *fn, size_t bytes) {
  if (bytes <= 65536) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                             static_cast<int>(bytes));
}

// split-mode (several problems per candidate) kernels exist for tile bounds up to 8
template <class Sh> constexpr bool split_shape() { return Sh::TT <= 8; }

template <class Sh, bool SPLIT> const void *run_dense_fn() {
  return reinterpret_cast<const void *>(&run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, SPLIT>);
}
template <class Sh, bool SPLIT> const void *batch_dense_fn() {
  return reinterpret_cast<const void *>(&batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, SPLIT>);
}

// F(integral_constant<bool, SPLIT>) for the requested form, hipErrorInvalidValue where the
// shape has no split-mode kernel
template <class Sh, class F> hipError_t with_split(bool split, F &&f) {
  if (!split) return f(std::false_type{});
  if constexpr (split_shape<Sh>()) return f(std::true_type{});
  return hipErrorInvalidValue;
}


template <class Sh> hipError_t run_dense_k(const RunParams &p, uint32_t blocks, hipStream_t s,
                                            bool split) {
  return with_split<Sh>(split, [&](auto sp) {
    constexpr bool SPLIT = decltype(sp)::value;
    const hipError_t e = dense_lds(run_dense_fn<Sh, SPLIT>(), p.wave_lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((run_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, SPLIT>), dim3(blocks),
                       dim3(64), p.wave_lds, s, p);
    return hipGetLastError();
  });
}

template <class Sh> hipError_t batch_dense_k(const BatchParams &bp, uint32_t blocks,
                                              hipStream_t s, bool split) {
  return with_split<Sh>(split, [&](auto sp) {
    constexpr bool SPLIT = decltype(sp)::value;
    hipLaunchKernelGGL((batch_dense_kernel<Sh::TT, Sh::G, Sh::W, Sh::D, SPLIT>), dim3(blocks),
                       dim3(64), bp.wave_lds, s, bp);
    return hipGetLastError();
  });
}

}  // namespace
}  // namespace sqsg
